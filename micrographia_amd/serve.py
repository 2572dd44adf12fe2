"""Serving: continuous batching of simulation requests on one GPU.

The reference has no serving layer (its nearest prior art is the page host's two batches in
flight, web/src/lib/host.ts:152-173).  Requests arrive over time; each is admitted into the batched
engine's device arena as soon as its slot, list and cell ranges fit in the free memory (the
admission rule is the device-memory budget), runs inside the shared launches with every other
admitted request, and leaves when its steps are done, freeing its ranges for waiting requests.
State stays on the device from admission to completion; only the final state is read back.
Each request's result is bit-identical to running it alone.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

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


class Server:
    """Continuous batching over a BatchEngine.

    budget_bytes: device memory for the arenas (the admission budget); chunk_s: target duration
    of one enqueued chunk of steps (the admission and retirement granularity)."""

    def __init__(self, budget_bytes: int, max_systems: int = 4096, chunk_s: float = 0.05, skin: float = 0.3,
                 workers: int = 8, keep_states: bool = False):
        self.batch = BatchEngine(budget_bytes=budget_bytes, max_systems=max_systems, graph_chunk=0)
        self.chunk_s = chunk_s
        self.skin = skin
        self.pool = cf.ThreadPoolExecutor(max_workers=workers)
        self.keep_states = keep_states

    def close(self) -> None:
        self.pool.shutdown(wait=True)
        self.batch.close()

    def serve(self, requests: List[Request], realtime: bool = True, on_done: Optional[Callable[[Completed], None]] = None) -> ServeReport:
        """Serve `requests` (arrival times in seconds from the start).  With realtime=False every
        request is treated as having arrived at time 0 (a closed batch)."""
        b = self.batch
        reqs = sorted(requests, key=lambda r: r.arrival)
        upcoming = deque(reqs)
        waiting: deque = deque()  # (request, future of (sys, state))
        active: Dict[int, tuple] = {}
        done: List[Completed] = []
        rejected: List[Request] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        iters = 16
        chunks = 0
        peak = 0
        atom_steps = 0
        busy = 0.0
        while upcoming or waiting or active:
            t = now()
            while upcoming and (not realtime or upcoming[0].arrival <= t):
                r = upcoming.popleft()
                if not realtime:
                    r = dataclasses.replace(r, arrival=0.0)
                waiting.append((r, self.pool.submit(r.build)))
            # admit in arrival order; a request that does not fit yet lets smaller ones pass
            still = deque()
            for r, fut in waiting:
                if not fut.done():
                    still.append((r, fut))
                    continue
                sys, st = fut.result()
                i = b.add(sys, st, r.steps, skin=self.skin)
                if i < 0:
                    if not active:
                        rejected.append(r)  # does not fit even in an empty arena
                    else:
                        still.append((r, fut))
                    continue
                active[i] = (r, now(), sys)
            waiting = still
            peak = max(peak, len(active))
            if not active:
                if upcoming and realtime:
                    time.sleep(max(0.0, min(0.01, upcoming[0].arrival - now())))
                elif waiting:
                    time.sleep(0.001)
                continue
            c0 = time.perf_counter()
            b.run(iters)
            left = b.poll()
            dt = time.perf_counter() - c0
            busy += dt
            chunks += 1
            # aim the next chunk at chunk_s
            per = dt / iters
            iters = int(max(4, min(1000, self.chunk_s / max(per, 1e-6))))
            for i, rem in left.items():
                if rem == 0 and i in active:
                    r, admitted, sys = active.pop(i)
                    st = b.read_state(i)
                    b.remove(i)
                    c = Completed(r, admitted, now(), st.hash(), st if self.keep_states else None)
                    atom_steps += r.n * r.steps
                    done.append(c)
                    if on_done:
                        on_done(c)
        return ServeReport(done, now(), atom_steps, chunks, peak, busy, rejected)


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
