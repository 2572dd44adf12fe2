"""Serving: continuous batching of simulation requests on one GPU.

The reference has no serving layer (its nearest prior art is the page host's two batches in
flight, web/src/lib/host.ts:152-173).  Requests arrive over time.  Small and medium requests are
admitted into the batched engine's device arena, shortest job first, as soon as their slot, list
and cell ranges fit in the free memory (the admission rule is the device-memory budget); they run
inside the shared launches with every other admitted request and leave when their steps are
done, freeing their ranges for waiting requests.  Requests large enough to fill the GPU alone run
in the single-system engine on a second lane, and so do requests whose systems the batched engine
does not run (compartments, membrane, piston, bonds).  State stays on the device from admission to
completion; only the final state is read back.  Each request's result is bit-identical to running
it alone.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

from ._lib import RH_ERR_UNSUPPORTED, RahmanError
from .batch import BatchEngine
from .workloads import Request


@dataclass
class Completed:
    request: Request
    admitted: float
    finished: float
    state_hash: int
    state: Optional[object] = None

    @property
    def latency(self) -> float:
        return self.finished - self.request.arrival


@dataclass
class ServeReport:
    completed: List[Completed]
    makespan: float
    atom_steps: int
    chunks: int
    peak_active: int
    busy_s: float = 0.0
    rejected: List[Request] = field(default_factory=list)

    def percentile(self, q: float) -> float:
        lat = sorted(c.latency for c in self.completed)
        if not lat:
            return float("nan")
        k = min(len(lat) - 1, max(0, int(round(q / 100.0 * (len(lat) - 1)))))
        return lat[k]

    def summary(self) -> Dict[str, float]:
        return {
            "requests": len(self.completed),
            "rejected": len(self.rejected),
            "makespan_s": self.makespan,
            "atom_steps": self.atom_steps,
            "atom_steps_per_s": self.atom_steps / self.makespan if self.makespan > 0 else 0.0,
            "p50_latency_s": self.percentile(50),
            "p99_latency_s": self.percentile(99),
            "max_latency_s": self.percentile(100),
            "chunks": self.chunks,
            "peak_active": self.peak_active,
        }


def _solo_mode(sys, n, skin):
    import micrographia_amd as rh

    return rh.verlet3_with_skin(sys, n, skin) if isinstance(sys, rh.System3) else rh.verlet_with_skin(sys, n, skin)


def _solo_bytes(sys, n, skin) -> int:
    from .engine import bytes_needed

    return bytes_needed(sys, n, _solo_mode(sys, n, skin))


def _solo_engine(sys, st, skin):
    import micrographia_amd as rh

    if isinstance(sys, rh.System3):
        return rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, st.n(), skin))
    return rh.GpuEngine(sys, st, rh.verlet_with_skin(sys, st.n(), skin))


class Server:
    """Continuous batching over a BatchEngine, with a second lane for large requests.

    - Requests with fewer than `big_n` particles go to the batch lane: they are admitted
      shortest job first (particles x steps) while their ranges fit the arena (the admission
      budget) and while the particles already active stay under `active_atoms` (enough to fill
      the GPU; more would only slow every active request down), and advance together.
    - Requests with `big_n` particles or more fill the GPU on their own: a worker thread runs
      them one at a time, shortest job first, in the single-system engine, concurrently with
      the batch lane (each lane has its own HIP stream).  Requests the batched engine refuses
      (systems it does not run) take this lane too; one the single-system engine also refuses
      is reported as rejected.

    Every result is bit-identical to running the request alone, whatever lane it took.
    budget_bytes: the device memory serving may use, shared by both lanes; big_bytes of it are
    the single-system lane's (one engine at a time; a request whose engine would need more is
    rejected), the rest is the batch arena (default: a quarter, at most 4 GiB, for that lane).
    chunk_s: target duration of one enqueued chunk of batch iterations (the admission and
    retirement granularity)."""

    def __init__(self, budget_bytes: int, max_systems: int = 4096, chunk_s: float = 0.02, skin: float = 0.3,
                 workers: int = 8, keep_states: bool = False, big_n: int = 131072, active_atoms: int = 6_000_000,
                 tuning=None, big_bytes: Optional[int] = None):
        self.big_bytes = int(big_bytes) if big_bytes is not None else min(int(budget_bytes) // 4, 4 << 30)
        if not 0 < self.big_bytes < budget_bytes:
            raise ValueError("big_bytes must leave part of the budget to the batch arena")
        self.budget_bytes = int(budget_bytes)
        self.batch = BatchEngine(budget_bytes=self.budget_bytes - self.big_bytes, max_systems=max_systems, graph_chunk=0)
        if tuning is not None:
            tuning.install(self.batch)  # an autotune.BatchTuner's per-bucket tile settings
        self.chunk_s = chunk_s
        self.skin = skin
        self.pool = cf.ThreadPoolExecutor(max_workers=workers)
        self.keep_states = keep_states
        self.big_n = big_n
        self.active_atoms = active_atoms

    def close(self) -> None:
        self.pool.shutdown(wait=True)
        self.batch.close()

    def serve(self, requests: List[Request], realtime: bool = True, on_done: Optional[Callable[[Completed], None]] = None) -> ServeReport:
        """Serve `requests` (arrival times in seconds from the start).  With realtime=False every
        request is treated as having arrived at time 0 (a closed batch)."""
        b = self.batch
        upcoming = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: List[tuple] = []   # batch lane: (request, future of (sys, state))
        active: Dict[int, tuple] = {}
        done: List[Completed] = []
        rejected: List[Request] = []
        lock = threading.Lock()
        big_cv = threading.Condition(lock)
        big_queue: List[tuple] = []
        big_open = [True]
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        stats = {"atom_steps": 0, "chunks": 0, "peak": 0, "busy": 0.0}

        def finish(c: Completed) -> None:
            with lock:
                done.append(c)
                stats["atom_steps"] += c.request.n * c.request.steps
            if on_done:
                on_done(c)

        def big_lane() -> None:
            while True:
                with big_cv:
                    while big_open[0] and not any(f.done() for _, f in big_queue):
                        big_cv.wait(0.005)
                    ready = [(r, f) for r, f in big_queue if f.done()]
                    if not ready:
                        if not big_open[0] and not big_queue:
                            return
                        big_cv.wait(0.005)  # inputs still being built
                        continue
                    r, f = min(ready, key=lambda rf: rf[0].n * rf[0].steps)
                    big_queue.remove((r, f))
                sys, st = f.result()
                admitted = now()
                try:
                    if _solo_bytes(sys, st.n(), self.skin) > self.big_bytes:
                        raise RahmanError(-3, "over the single-system lane's share of the budget")
                    e = _solo_engine(sys, st, self.skin)
                except RahmanError:
                    with lock:
                        rejected.append(r)  # over the lane's budget, or the engine refuses it
                    continue
                try:
                    e.step(r.steps)
                    fin = e.read_state()
                finally:
                    e.close()
                finish(Completed(r, admitted, now(), fin.hash(), fin if self.keep_states else None))

        worker = threading.Thread(target=big_lane, name="big-lane", daemon=True)
        worker.start()
        iters = 16
        active_n = 0
        try:
            while upcoming or waiting or active:
                t = now()
                while upcoming and (not realtime or upcoming[0].arrival <= t):
                    r = upcoming.popleft()
                    if not realtime:
                        r = dataclasses.replace(r, arrival=0.0)
                    fut = self.pool.submit(r.build)
                    if r.n >= self.big_n:
                        with big_cv:
                            big_queue.append((r, fut))
                            big_cv.notify()
                    else:
                        waiting.append((r, fut))
                # batch lane admission: shortest job first, within the arena and the active cap
                waiting.sort(key=lambda rf: rf[0].n * rf[0].steps)
                still = []
                for r, fut in waiting:
                    if not fut.done() or (active and active_n + r.n > self.active_atoms):
                        still.append((r, fut))
                        continue
                    sys, st = fut.result()
                    try:
                        i = b.add(sys, st, r.steps, skin=self.skin)
                    except RahmanError as err:
                        if err.code != RH_ERR_UNSUPPORTED:
                            raise  # a real failure, not a refusal
                        # a system the batched engine does not run (compartments, membrane,
                        # piston, bonds): the single-system lane takes it
                        with big_cv:
                            big_queue.append((r, fut))
                            big_cv.notify()
                        continue
                    if i < 0:
                        if not active:
                            rejected.append(r)  # does not fit even in an empty arena
                        else:
                            still.append((r, fut))
                        continue
                    active[i] = (r, now())
                    active_n += r.n
                waiting = still
                stats["peak"] = max(stats["peak"], len(active))
                if not active:
                    if upcoming and realtime and not waiting:
                        time.sleep(max(0.0, min(0.01, upcoming[0].arrival - now())))
                    elif waiting:
                        time.sleep(0.001)
                    continue
                c0 = time.perf_counter()
                b.run(iters)
                left = b.poll()
                dt = time.perf_counter() - c0
                stats["busy"] += dt
                stats["chunks"] += 1
                iters = int(max(4, min(1000, self.chunk_s / max(dt / iters, 1e-6))))
                for i, rem in left.items():
                    if rem == 0 and i in active:
                        r, admitted = active.pop(i)
                        active_n -= r.n
                        st = b.read_state(i)
                        b.remove(i)
                        finish(Completed(r, admitted, now(), st.hash(), st if self.keep_states else None))
        finally:
            with big_cv:
                big_open[0] = False
                big_cv.notify()
            worker.join()
        return ServeReport(done, now(), stats["atom_steps"], stats["chunks"], stats["peak"], stats["busy"], rejected)


def serve_plain(requests: List[Request], skin: float = 0.3, realtime: bool = True) -> ServeReport:
    """The same requests served one at a time by the single-system engines (FIFO), for
    comparison: one system per launch."""
    import micrographia_amd as rh

    reqs = sorted(requests, key=lambda r: r.arrival)
    done: List[Completed] = []
    t0 = time.perf_counter()
    atom_steps = 0
    for r in reqs:
        if realtime:
            while time.perf_counter() - t0 < r.arrival:
                time.sleep(0.001)
        sys, st = r.build()
        admitted = time.perf_counter() - t0
        if isinstance(sys, rh.System3):
            e = rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, st.n(), skin))
        else:
            e = rh.GpuEngine(sys, st, rh.verlet_with_skin(sys, st.n(), skin))
        e.step(r.steps)
        fin = e.read_state()
        e.close()
        atom_steps += r.n * r.steps
        r2 = r if realtime else dataclasses.replace(r, arrival=0.0)
        done.append(Completed(r2, admitted, time.perf_counter() - t0, fin.hash()))
    return ServeReport(done, time.perf_counter() - t0, atom_steps, len(reqs), 1)
