"""The serving loop's host logic on the CPU, with stand-ins for the device engines: every request
completes exactly once on the lane its size selects, the batch lane admits shortest job first
within the arena and the active-particle cap, requests that can never fit are rejected, and the
report's percentiles are computed over completion latencies."""
import dataclasses

import pytest

from micrographia_amd import serve
from micrographia_amd.workloads import Request


class FakeState:
    def __init__(self, n, tag):
        self._n, self.tag = n, tag

    def n(self):
        return self._n

    def hash(self):
        return hash(("done", self.tag))


class FakeRequest(Request):
    def build(self):
        return ("sys", self.index), FakeState(self.n, self.index)


class FakeBatch:
    """Admits while the particle budget allows; every member advances one step per iteration."""

    log = []

    def __init__(self, budget_bytes, max_systems, graph_chunk):
        self.capacity_atoms = budget_bytes
        self.members = {}
        self.next_id = 0

    def add(self, sys, st, steps, skin=0.3):
        used = sum(m["n"] for m in self.members.values())
        if used + st.n() > self.capacity_atoms:
            return -1
        i = self.next_id
        self.next_id += 1
        self.members[i] = {"n": st.n(), "left": steps, "tag": st.tag}
        FakeBatch.log.append(("add", st.tag))
        return i

    def run(self, iters):
        for m in self.members.values():
            m["left"] = max(0, m["left"] - iters)

    def poll(self):
        return {i: m["left"] for i, m in self.members.items()}

    def read_state(self, i):
        return FakeState(self.members[i]["n"], self.members[i]["tag"])

    def remove(self, i):
        del self.members[i]

    def close(self):
        pass


class FakeSolo:
    ran = []

    def __init__(self, sys, st):
        self.st = st

    def step(self, steps):
        FakeSolo.ran.append(self.st.tag)

    def read_state(self):
        return FakeState(self.st.n(), self.st.tag)

    def close(self):
        pass


@pytest.fixture
def fakes(monkeypatch):
    FakeBatch.log = []
    FakeSolo.ran = []
    monkeypatch.setattr(serve, "BatchEngine", FakeBatch)
    monkeypatch.setattr(serve, "_solo_engine", lambda sys, st, skin: FakeSolo(sys, st))
    monkeypatch.setattr(serve, "_solo_bytes", lambda sys, n, skin: n)  # a particle a byte


def reqs(spec):
    return [FakeRequest(k, 3, n, 0.8, 1.0, k + 1, steps, 0.0) for k, (n, steps) in enumerate(spec)]


def test_every_request_completes_once_on_its_lane(fakes):
    rs = reqs([(100, 50), (5000, 10), (300, 20), (9000, 5), (200, 80)])
    srv = serve.Server(budget_bytes=20_000, big_bytes=10_000, chunk_s=0.001, big_n=4000)
    rep = srv.serve(rs, realtime=False)
    srv.close()
    assert sorted(c.request.index for c in rep.completed) == [0, 1, 2, 3, 4]
    assert sorted(FakeSolo.ran) == [1, 3]                      # n >= big_n: single-system lane
    assert sorted(t for _, t in FakeBatch.log) == [0, 2, 4]    # the rest: batch lane
    for c in rep.completed:
        assert c.state_hash == hash(("done", c.request.index))


def test_batch_lane_is_shortest_job_first_under_the_active_cap(fakes):
    # work = n * steps: request 2 (100 x 10) < 0 (200 x 20) < 1 (150 x 100)
    rs = reqs([(200, 20), (150, 100), (100, 10)])
    srv = serve.Server(budget_bytes=20_000, big_bytes=10_000, chunk_s=0.001, big_n=10_000, active_atoms=150)
    rep = srv.serve(rs, realtime=False)
    srv.close()
    assert [t for _, t in FakeBatch.log] == [2, 0, 1]
    assert len(rep.completed) == 3


def test_requests_that_never_fit_are_rejected(fakes):
    rs = reqs([(100, 10), (50_000, 10)])
    srv = serve.Server(budget_bytes=2_000, big_bytes=1_000, chunk_s=0.001, big_n=10**9)
    rep = srv.serve(rs, realtime=False)
    srv.close()
    assert [r.index for r in rep.rejected] == [1]
    assert [c.request.index for c in rep.completed] == [0]


def test_the_single_system_lane_counts_against_the_budget(fakes):
    """Both lanes share budget_bytes: the batch arena gets what the single-system lane's share
    leaves, and a large request whose engine would need more than that share is rejected."""
    rs = reqs([(5000, 10), (12_000, 10), (300, 20)])
    srv = serve.Server(budget_bytes=30_000, big_bytes=10_000, chunk_s=0.001, big_n=4000)
    assert srv.batch.capacity_atoms == 20_000
    rep = srv.serve(rs, realtime=False)
    srv.close()
    assert [r.index for r in rep.rejected] == [1]
    assert sorted(c.request.index for c in rep.completed) == [0, 2]
    with pytest.raises(ValueError):
        serve.Server(budget_bytes=1_000, big_bytes=1_000)


def test_thousands_of_admissions_and_retirements(fakes, monkeypatch):
    """Churn: 2,000 requests of sizes over three decades through a small arena; every one
    completes exactly once, on the lane its size selects, with the arena never over-full."""
    import random

    rng = random.Random(5)
    spec = [(int(10 ** rng.uniform(1.5, 3.7)), rng.randint(1, 60)) for _ in range(2000)]
    rs = reqs(spec)
    peak = [0]

    class Watched(FakeBatch):
        def add(self, sys, st, steps, skin=0.3):
            i = super().add(sys, st, steps, skin)
            peak[0] = max(peak[0], sum(m["n"] for m in self.members.values()))
            return i

    monkeypatch.setattr(serve, "BatchEngine", Watched)
    srv = serve.Server(budget_bytes=40_000, big_bytes=8_000, chunk_s=0.0005, big_n=4000)
    rep = srv.serve(rs, realtime=False)
    srv.close()
    assert sorted(c.request.index for c in rep.completed) == list(range(2000)) and not rep.rejected
    assert sorted(FakeSolo.ran) == sorted(k for k, (n, _) in enumerate(spec) if n >= 4000)
    assert peak[0] <= 32_000 and rep.peak_active > 10


def test_percentiles():
    r = FakeRequest(0, 3, 10, 0.8, 1.0, 1, 10, 0.0)
    done = [serve.Completed(dataclasses.replace(r, index=k), 0.0, float(k + 1), 0) for k in range(100)]
    rep = serve.ServeReport(done, 100.0, 0, 0, 0)
    assert rep.percentile(50) == 51.0 or rep.percentile(50) == 50.0
    assert rep.percentile(99) == 99.0 or rep.percentile(99) == 100.0
    assert rep.percentile(100) == 100.0
