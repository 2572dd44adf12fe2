"""BatchEngine: many independent systems advancing together on one GPU.

The reference has no counterpart (it runs one system per engine; its 2D compartments,
crates/md-core/src/system.rs:41-48, are the nearest prior art).  Systems of any size, 2D and 3D
mixed, share one device arena; every kernel of a step is one launch for all of them, and chunks
of steps are replayed from a captured hipGraph.  Systems can join and leave between runs, which
is what the serving loop (micrographia_amd.serve) builds on.  Each system's trajectory is
bit-identical to running it alone in GpuEngine3 / GpuEngine.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Union

import numpy as np

from ._lib import RH_ERR_NOMEM, CLedger3, CListStats, RahmanError, check, load, ptr
from .core import State, State3, System, System3
from .engine import Ledger3, ListStats, _stats


class BatchEngine:
    def __init__(self, budget_bytes: int = 0, max_systems: int = 4096, graph_chunk: int = 100):
        h = C.c_void_p()
        check(load().rh_batch_create(int(budget_bytes), int(max_systems), C.byref(h)))
        self._h = h
        self.max_systems = int(max_systems)
        self._members: Dict[int, Union[System, System3]] = {}
        self._n: Dict[int, int] = {}
        self.set_graph_chunk(graph_chunk)

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().rh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, sys: Union[System, System3], state: Union[State, State3], steps: int, skin: float = 0.3) -> int:
        """Admit a system with `steps` steps to run; returns its id, or -1 if it does not fit now."""
        cs = sys.to_c()
        out = C.c_int32(-1)
        fn = load().rh_batch_add3 if isinstance(sys, System3) else load().rh_batch_add2
        check(fn(self._h, C.byref(cs), state.n(), ptr(state.pos), ptr(state.vel), ptr(state.species), float(skin), int(steps), C.byref(out)))
        if out.value >= 0:
            self._members[out.value] = sys
            self._n[out.value] = state.n()
        return int(out.value)

    def bytes_needed(self, sys: Union[System, System3], n: int, skin: float = 0.3) -> int:
        cs = sys.to_c()
        out = C.c_uint64()
        fn = load().rh_batch_bytes_needed3 if isinstance(sys, System3) else load().rh_batch_bytes_needed2
        check(fn(self._h, C.byref(cs), int(n), float(skin), C.byref(out)))
        return int(out.value)

    def remove(self, id_: int) -> None:
        check(load().rh_batch_remove(self._h, int(id_)))
        self._members.pop(id_, None)
        self._n.pop(id_, None)

    def set_steps(self, id_: int, steps: int) -> None:
        check(load().rh_batch_set_steps(self._h, int(id_), int(steps)))

    def run(self, iterations: int) -> None:
        """Enqueue `iterations` steps for every unfinished system (returns without waiting)."""
        check(load().rh_batch_run(self._h, int(iterations)))

    def sync(self) -> None:
        check(load().rh_batch_sync(self._h))

    def poll(self) -> Dict[int, int]:
        """Steps left per member (synchronises with the device)."""
        rem = np.zeros(self.max_systems, np.uint64)
        check(load().rh_batch_poll(self._h, ptr(rem), self.max_systems))
        return {i: int(rem[i]) for i in self._members}

    def step_all(self, steps: Optional[int] = None) -> None:
        """Advance every member by `steps` steps, or to the end of its own when fewer are left
        (all of them with steps=None), and wait.  A system whose list overflowed stops (its steps
        are gated) until poll() has regrown the list; the steps it lost are then run for it alone
        (the others paused at their targets).  Raises if a catch-up round makes no progress: an
        overflowed list that cannot grow in the arena as it stands."""
        before = self.poll()
        if steps is None:
            steps = max(before.values(), default=0)
        target = {i: max(0, v - int(steps)) for i, v in before.items()}
        self.run(steps)
        left = self.poll()
        while any(left[i] > target[i] for i in left):
            deficit = {i: left[i] - target[i] for i in left}
            for i, d in deficit.items():
                self.set_steps(i, d)  # lagging members run what they lost, the rest pause
            self.run(max(deficit.values()))
            now = self.poll()
            for i in left:
                self.set_steps(i, target[i] + now[i])
            if not any(now[i] < deficit[i] for i in deficit if deficit[i]):
                stuck = sorted(i for i, d in deficit.items() if d)
                raise RahmanError(RH_ERR_NOMEM, f"systems {stuck} cannot continue: their lists overflowed and the arena has no room to grow them")
            left = self.poll()

    def read_state(self, id_: int) -> Union[State, State3]:
        sys = self._members[id_]
        n = self._n[id_]
        d = 3 if isinstance(sys, System3) else 2
        pos = np.zeros((n, d), np.int64)
        vel = np.zeros((n, d), np.int32)
        sp = np.zeros(n, np.uint32)
        check(load().rh_batch_read_state(self._h, int(id_), ptr(pos), ptr(vel), ptr(sp)))
        return State3(pos, vel, sp) if d == 3 else State(pos, vel, sp)

    def read_ledger(self, id_: int) -> Ledger3:
        """The system's ledger since it joined: energy, impulse and hits per wall (2 axis +
        side) and for the bath (entry 6), as GpuEngine3::read_ledger (engine3.rs:645-647) and
        GpuEngine's for 2D (entries 0-3 the walls)."""
        l = CLedger3()
        check(load().rh_batch_read_ledger(self._h, int(id_), C.byref(l)))
        return Ledger3(np.array(l.energy[:]), np.array(l.impulse[:]), np.array(l.hits[:], dtype=np.uint64))

    def list_stats(self, id_: int) -> ListStats:
        s = CListStats()
        check(load().rh_batch_list_stats(self._h, int(id_), C.byref(s)))
        return _stats(s)

    def set_graph_chunk(self, chunk: int) -> None:
        check(load().rh_batch_set_graph_chunk(self._h, int(chunk)))

    def set_tuning(self, dim: int, size_class: int, variant: int, cell_factor: float, block: int = 256) -> None:
        """Tile settings (list variant: 1 half list, built in 3D over the row-ordered copy where
        the box allows, 2 the half list over the cell stencil, 0 every pair at both ends; cell
        occupancy; half-list block of 256, 512 or 1024 slots) for systems of this dimension and
        size class (floor(log2 n)) added from now on; trajectories do not depend on them."""
        check(load().rh_batch_set_tuning(self._h, int(dim), int(size_class), int(variant), float(cell_factor), int(block)))

    def capacity(self):
        s, l, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        a = C.c_uint32()
        check(load().rh_batch_capacity(self._h, C.byref(s), C.byref(l), C.byref(c), C.byref(a)))
        return {"slots": int(s.value), "list_entries": int(l.value), "cells": int(c.value), "active": int(a.value)}

    def members(self) -> List[int]:
        return sorted(self._members)

    def stream(self) -> int:
        v = C.c_void_p()
        check(load().rh_batch_stream(self._h, C.byref(v)))
        return int(v.value or 0)
