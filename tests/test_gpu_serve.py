"""The serving loop: requests arriving over time are admitted against the device-memory budget,
batched with whatever else is running, and each result is bit-identical to the request run
alone in the single-system engine."""
from dataclasses import dataclass

import pytest

import micrographia_amd as rh
from micrographia_amd.serve import Server, serve_plain
from micrographia_amd.workloads import Request, request_stream

pytestmark = pytest.mark.gpu


def solo_hash(r, skin=0.3):
    sys, st = r.build()
    if isinstance(sys, rh.System3):
        e = rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, st.n(), skin))
    else:
        e = rh.GpuEngine(sys, st, rh.verlet_with_skin(sys, st.n(), skin))
    e.step(r.steps)
    h = e.read_state().hash()
    e.close()
    return h


def test_served_results_equal_solo_runs():
    reqs = request_stream(400.0, 0.05, meta_seed=5, n_lo=300, n_hi=30000, steps_lo=20, steps_hi=400)
    assert len(reqs) >= 8
    srv = Server(budget_bytes=1 << 30, max_systems=64, chunk_s=0.01)
    rep = srv.serve(reqs, realtime=True)
    srv.close()
    assert len(rep.completed) == len(reqs) and not rep.rejected
    assert rep.peak_active >= 2  # requests really were batched together
    for c in rep.completed:
        assert c.state_hash == solo_hash(c.request), f"request {c.request.index} (n={c.request.n}, dim={c.request.dim})"
        assert c.latency >= 0 and c.finished >= c.admitted


def test_small_budget_queues_and_stays_exact():
    reqs = request_stream(1000.0, 0.02, meta_seed=8, n_lo=2000, n_hi=20000, steps_lo=30, steps_hi=120)
    srv = Server(budget_bytes=96 << 20, max_systems=16, chunk_s=0.005)
    rep = srv.serve(reqs, realtime=False)
    srv.close()
    done = {c.request.index: c for c in rep.completed}
    assert len(done) + len(rep.rejected) == len(reqs) and len(done) >= 2
    for c in rep.completed:
        assert c.state_hash == solo_hash(c.request)
    plain = serve_plain([c.request for c in rep.completed], realtime=False)
    assert sorted(c.state_hash for c in plain.completed) == sorted(c.state_hash for c in rep.completed)


def test_both_lanes_together_stay_exact():
    """Requests above big_n run on the single-system lane while the rest are batched; both
    lanes share the GPU at once."""
    reqs = request_stream(600.0, 0.03, meta_seed=21, n_lo=500, n_hi=40000, steps_lo=40, steps_hi=200)
    srv = Server(budget_bytes=1 << 30, max_systems=64, chunk_s=0.005, big_n=8000)
    rep = srv.serve(reqs, realtime=True)
    srv.close()
    assert len(rep.completed) == len(reqs)
    assert any(r.n >= 8000 for r in reqs) and any(r.n < 8000 for r in reqs)
    for c in rep.completed:
        assert c.state_hash == solo_hash(c.request), f"request {c.request.index} (n={c.request.n})"


@dataclass
class ThermostatRequest(Request):
    """A request whose system has a thermostat: the batched engine runs it in 3D and refuses it
    in 2D (the thermostat region belongs to the 2D boundary path)."""

    def build(self):
        sys, st = super().build()
        return sys.with_thermostat(1.2, 2.0), st


def test_requests_the_batch_refuses_take_the_single_system_lane():
    reqs = request_stream(600.0, 0.03, meta_seed=31, n_lo=500, n_hi=6000, steps_lo=40, steps_hi=150)
    reqs = [ThermostatRequest(**{f: getattr(r, f) for f in r.__dataclass_fields__}) if k % 2 == 0 else r for k, r in enumerate(reqs)]
    assert any(isinstance(r, ThermostatRequest) and r.dim == 2 for r in reqs)
    assert any(isinstance(r, ThermostatRequest) and r.dim == 3 for r in reqs)
    srv = Server(budget_bytes=1 << 30, max_systems=64, chunk_s=0.005, big_n=1 << 30)
    rep = srv.serve(reqs, realtime=False)
    srv.close()
    assert len(rep.completed) == len(reqs) and not rep.rejected
    for c in rep.completed:
        assert c.state_hash == solo_hash(c.request), f"request {c.request.index} (n={c.request.n}, dim={c.request.dim})"
