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


def test_serving_at_full_request_sizes_against_the_oracle():
    """The request sizes of the serving workload's extremes in one stream: two N = 1,000,000
    requests (3D and 2D; they take the single-system lane), a 2D request of 5,000 steps, and
    small 2D and 3D requests arriving while those run, at densities and temperatures drawn
    from its ranges.  The small and the long requests equal the oracle's trajectories bit for
    bit; the large ones equal their solo runs."""
    import concurrent.futures as cf

    import numpy as np

    from oracle_twin import oracle_state

    rng = np.random.default_rng(11)
    reqs = []
    for k in range(10):
        reqs.append(Request(k, 2 + k % 2, int(rng.integers(500, 3001)), float(rng.uniform(0.3, 0.9)), float(rng.uniform(0.5, 2.0)),
                            1000 + k, int(rng.integers(100, 301)), 0.005 * k))
    reqs.append(Request(10, 2, 1000, 0.7, 1.0, 2001, 5000, 0.0))
    reqs.append(Request(11, 3, 1_000_000, 0.8, 1.2, 2002, 100, 0.01))
    reqs.append(Request(12, 2, 1_000_000, 0.6, 1.0, 2003, 100, 0.02))
    srv = Server(budget_bytes=16 << 30, max_systems=64, chunk_s=0.01, keep_states=True)
    rep = srv.serve(reqs, realtime=True)
    srv.close()
    assert len(rep.completed) == len(reqs) and not rep.rejected
    done = {c.request.index: c for c in rep.completed}
    checked = [r for r in reqs if r.n < 1_000_000]
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        want = list(ex.map(lambda r: oracle_state(*r.build(), r.steps), checked))
    for r, w in zip(checked, want):
        assert done[r.index].state == w, f"request {r.index} (n={r.n}, dim={r.dim}, steps={r.steps}) differs from the oracle"
    for r in reqs:
        if r.n >= 1_000_000:
            assert done[r.index].state_hash == solo_hash(r), f"request {r.index} (n={r.n}, dim={r.dim})"


@dataclass
class BoundaryRequest(Request):
    """A request whose system carries a boundary: in 3D the Andersen bath (the batched engine
    runs it), in 2D a harmonic bond between its first two particles (neighbours on the start
    lattice; the batched engine refuses bonds)."""

    def build(self):
        sys, st = super().build()
        if self.dim == 2:
            return sys.with_bonds([rh.BondType.harmonic(40.0, 1.0)], [(0, 1, 0)]), st
        return sys.with_thermostat(1.2, 2.0), st


def test_requests_the_batch_refuses_take_the_single_system_lane():
    reqs = request_stream(600.0, 0.03, meta_seed=31, n_lo=500, n_hi=6000, steps_lo=40, steps_hi=150)
    reqs = [BoundaryRequest(**{f: getattr(r, f) for f in r.__dataclass_fields__}) if k % 2 == 0 else r for k, r in enumerate(reqs)]
    assert any(isinstance(r, BoundaryRequest) and r.dim == 2 for r in reqs)
    assert any(isinstance(r, BoundaryRequest) and r.dim == 3 for r in reqs)
    srv = Server(budget_bytes=1 << 30, max_systems=64, chunk_s=0.005, big_n=1 << 30)
    rep = srv.serve(reqs, realtime=False)
    srv.close()
    assert len(rep.completed) == len(reqs) and not rep.rejected
    for c in rep.completed:
        assert c.state_hash == solo_hash(c.request), f"request {c.request.index} (n={c.request.n}, dim={c.request.dim})"
