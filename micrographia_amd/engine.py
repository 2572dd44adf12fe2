"""GpuEngine3 / GpuEngine: the MI355X engines behind the reference's engine API.

Method names and meanings follow md::GpuEngine3 (crates/md/src/engine3.rs:307-648) and
md::GpuEngine (crates/md/src/engine.rs:567-1098).  The reference's readbacks are futures; here
they return once the device has answered.  Errors raise RahmanError (the reference panics).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check, load, ptr
from .core import ForceMode, State, State3, System, System3, verlet3, verlet_for


@dataclass
class Ledger3:
    energy: np.ndarray
    impulse: np.ndarray
    hits: np.ndarray


@dataclass
class Ledger:
    """md_core::walls::Ledger (walls.rs:41-57): per entry (walls 0-3 = 2 axis + side, membrane
    faces 4 and 5, thermostat 6) the energy and impulse exchanged and the hits."""

    energy: np.ndarray
    impulse: np.ndarray
    hits: np.ndarray


@dataclass
class CompartmentLedger:
    """md_core::walls::CompartmentLedger: the ledger per entry and compartment along the wall,
    arrays of shape (7, per_wall)."""

    energy: np.ndarray
    impulse: np.ndarray
    hits: np.ndarray

    def totals(self) -> Ledger:
        return Ledger(self.energy.sum(axis=1), self.impulse.sum(axis=1), self.hits.sum(axis=1))


@dataclass
class ListStats:
    rebuilds: int
    overflowed: bool
    cap: int
    max_count: int
    steps_done: int
    launches_skipped: int


def _stats(s: _lib.CListStats) -> ListStats:
    return ListStats(int(s.rebuilds), bool(s.overflowed), int(s.cap), int(s.max_count), int(s.steps_done), int(s.launches_skipped))


def bytes_needed(sys, n: int, mode: Optional[ForceMode] = None) -> int:
    """Device memory an engine for `n` particles of `sys` would allocate (a plan on the host; no
    device work).  The serving loop counts its single-system lane against its budget with it."""
    cs = sys.to_c()
    m = mode.to_c() if mode is not None else None
    out = C.c_uint64()
    fn = load().rh_engine3_bytes_needed if isinstance(sys, System3) else load().rh_engine2_bytes_needed
    check(fn(C.byref(cs), int(n), C.byref(m) if m is not None else None, C.byref(out)))
    return int(out.value)


class GpuEngine3:
    """md::GpuEngine3 on one MI355X."""

    def __init__(self, sys: System3, state: State3, mode: Optional[ForceMode] = None):
        self.sys = sys
        self.mode = mode if mode is not None else verlet3(sys, state.n())
        self._n = state.n()
        h = C.c_void_p()
        cs = sys.to_c()
        cm = self.mode.to_c()
        check(load().rh_engine3_create(C.byref(cs), self._n, ptr(state.pos), ptr(state.vel), ptr(state.species), C.byref(cm), C.byref(h)))
        self._h = h
        self._sort_interval = 100

    @staticmethod
    def new(sys: System3, state: State3) -> "GpuEngine3":
        """GpuEngine3::new: Verlet lists with a 0.4 sigma skin (engine3.rs:337-340)."""
        return GpuEngine3(sys, state)

    @staticmethod
    def with_mode(sys: System3, state: State3, mode: ForceMode) -> "GpuEngine3":
        return GpuEngine3(sys, state, mode)

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().rh_engine3_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n(self) -> int:
        return self._n

    @property
    def steps(self) -> int:
        v = C.c_int64()
        check(load().rh_engine3_steps(self._h, C.byref(v)))
        return int(v.value)

    @property
    def sort_interval(self) -> int:
        return self._sort_interval

    @sort_interval.setter
    def sort_interval(self, v: int) -> None:
        check(load().rh_engine3_set_sort_interval(self._h, int(v)))
        self._sort_interval = int(v)

    def step(self, steps: int) -> None:
        check(load().rh_engine3_step(self._h, int(steps)))

    def reverse(self) -> None:
        check(load().rh_engine3_reverse(self._h))

    def update_system(self, sys: System3) -> None:
        cs = sys.to_c()
        check(load().rh_engine3_update_system(self._h, C.byref(cs)))
        self.sys = sys

    def kick_region(self, lo: Sequence[float], hi: Sequence[float], scale: float, add: Sequence[float]) -> None:
        lo_a, hi_a, add_a = (np.ascontiguousarray(x, np.float64) for x in (lo, hi, add))
        check(load().rh_engine3_kick_region(self._h, ptr(lo_a), ptr(hi_a), float(scale), ptr(add_a)))

    def set_species(self, ids: Sequence[int], species: int) -> None:
        a = np.ascontiguousarray(ids, np.uint32)
        check(load().rh_engine3_set_species(self._h, ptr(a), len(a), int(species)))

    def set_species_velocity(self, species: int, v: Sequence[float]) -> None:
        a = np.ascontiguousarray(v, np.float64)
        check(load().rh_engine3_set_species_velocity(self._h, int(species), ptr(a)))

    def upload(self, state: State3) -> None:
        if state.n() != self._n:
            raise ValueError("upload needs the same particle count")
        check(load().rh_engine3_upload(self._h, ptr(state.pos), ptr(state.vel), ptr(state.species)))

    def read_state(self) -> State3:
        pos = np.zeros((self._n, 3), np.int64)
        vel = np.zeros((self._n, 3), np.int32)
        sp = np.zeros(self._n, np.uint32)
        check(load().rh_engine3_read_state(self._h, ptr(pos), ptr(vel), ptr(sp)))
        return State3(pos, vel, sp)

    def kicks(self) -> Tuple[np.ndarray, float]:
        k = np.zeros((self._n, 3), np.int32)
        pe = C.c_double()
        check(load().rh_engine3_kicks(self._h, ptr(k), C.byref(pe)))
        return k, float(pe.value)

    def device_bytes(self) -> int:
        """Device memory the engine holds."""
        v = C.c_uint64()
        check(load().rh_engine3_device_bytes(self._h, C.byref(v)))
        return int(v.value)

    def measure_kinetic(self) -> None:
        """Every particle's kinetic energy and their sum, from the half-step velocities
        (engine3.rs:618-624)."""
        check(load().rh_engine3_measure_kinetic(self._h))

    def read_kinetic(self) -> Tuple[np.ndarray, np.ndarray]:
        """(kinetic energies as f32, species), both in particle-id order; measures now
        (engine3.rs:626-631)."""
        ke = np.zeros(self._n, np.float32)
        sp = np.zeros(self._n, np.uint32)
        check(load().rh_engine3_read_kinetic(self._h, ptr(ke), ptr(sp)))
        return ke, sp

    def read_kinetic_total(self) -> float:
        v = C.c_double()
        check(load().rh_engine3_read_kinetic_total(self._h, C.byref(v)))
        return float(v.value)

    def read_potential(self) -> float:
        v = C.c_double()
        check(load().rh_engine3_read_potential(self._h, C.byref(v)))
        return float(v.value)

    def read_ledger(self) -> Ledger3:
        l = _lib.CLedger3()
        check(load().rh_engine3_read_ledger(self._h, C.byref(l)))
        return Ledger3(np.array(l.energy[:]), np.array(l.impulse[:]), np.array(l.hits[:], dtype=np.uint64))

    def read_list_stats(self) -> ListStats:
        s = _lib.CListStats()
        check(load().rh_engine3_list_stats(self._h, C.byref(s)))
        return _stats(s)

    def tune(self, variant: int, block: int = 512) -> None:
        """Kernel variant (0 full list; 1 half list in blocks of `block`: 256, 512, 1024, 2048 or
        4096, built from the row-ordered copy where the geometry allows; 2 the same half list
        built over the cell stencil); trajectories are identical."""
        check(load().rh_engine3_tune(self._h, int(variant), int(block)))

    def stream(self) -> int:
        v = C.c_void_p()
        check(load().rh_engine3_stream(self._h, C.byref(v)))
        return int(v.value or 0)

    def set_timing(self, on: bool) -> None:
        check(load().rh_engine3_set_timing(self._h, 1 if on else 0))

    def kernel_ms(self):
        f, r = C.c_double(), C.c_double()
        fl, rl = C.c_uint64(), C.c_uint64()
        check(load().rh_engine3_kernel_ms(self._h, C.byref(f), C.byref(fl), C.byref(r), C.byref(rl)))
        return float(f.value), int(fl.value), float(r.value), int(rl.value)


class GpuEngine:
    """md::GpuEngine (2D): f32 pair forces (or the all-integer tables) with md_core::System's
    boundaries: walls (hard, soft, heat bath), compartments, membrane, thermostat / fan / bath
    layer, piston, bonds, gravity and pinned species."""

    def __init__(self, sys: System, state: State, mode: Optional[ForceMode] = None):
        self.sys = sys
        self.mode = mode if mode is not None else verlet_for(sys, state.n())
        self._n = state.n()
        h = C.c_void_p()
        cs = sys.to_c()
        cm = self.mode.to_c()
        check(load().rh_engine2_create(C.byref(cs), self._n, ptr(state.pos), ptr(state.vel), ptr(state.species), C.byref(cm), C.byref(h)))
        self._h = h

    @staticmethod
    def new(sys: System, state: State) -> "GpuEngine":
        return GpuEngine(sys, state)

    @staticmethod
    def with_mode(sys: System, state: State, mode: ForceMode) -> "GpuEngine":
        return GpuEngine(sys, state, mode)

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().rh_engine2_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n(self) -> int:
        return self._n

    def step(self, steps: int) -> None:
        check(load().rh_engine2_step(self._h, int(steps)))

    def reverse(self) -> None:
        check(load().rh_engine2_reverse(self._h))

    def upload(self, state: State) -> None:
        check(load().rh_engine2_upload(self._h, ptr(state.pos), ptr(state.vel), ptr(state.species)))

    def read_state(self) -> State:
        pos = np.zeros((self._n, 2), np.int64)
        vel = np.zeros((self._n, 2), np.int32)
        sp = np.zeros(self._n, np.uint32)
        check(load().rh_engine2_read_state(self._h, ptr(pos), ptr(vel), ptr(sp)))
        return State(pos, vel, sp)

    def kicks(self) -> Tuple[np.ndarray, float]:
        k = np.zeros((self._n, 2), np.int32)
        pe = C.c_double()
        check(load().rh_engine2_kicks(self._h, ptr(k), C.byref(pe)))
        return k, float(pe.value)

    def device_bytes(self) -> int:
        """Device memory the engine holds."""
        v = C.c_uint64()
        check(load().rh_engine2_device_bytes(self._h, C.byref(v)))
        return int(v.value)

    def compute_forces(self) -> None:
        """A force evaluation at the current positions without moving (engine.rs:875-878)."""
        check(load().rh_engine2_compute_forces(self._h))

    def read_kinetic_total(self) -> float:
        v = C.c_double()
        check(load().rh_engine2_read_kinetic_total(self._h, C.byref(v)))
        return float(v.value)

    def read_list_stats(self) -> ListStats:
        s = _lib.CListStats()
        check(load().rh_engine2_list_stats(self._h, C.byref(s)))
        return _stats(s)

    def tune(self, variant: int, block: int = 512) -> None:
        check(load().rh_engine2_tune(self._h, int(variant), int(block)))

    def set_arith(self, arith: int) -> None:
        """md::Arith: 0 f32 pair forces, 1 all-integer tables (pair forces only; bit-identical on
        every device, md_ref::IntEngine's integer forces)."""
        check(load().rh_engine2_set_arith(self._h, int(arith)))

    def stream(self) -> int:
        v = C.c_void_p()
        check(load().rh_engine2_stream(self._h, C.byref(v)))
        return int(v.value or 0)

    @property
    def steps(self) -> int:
        v = C.c_int64()
        check(load().rh_engine2_steps(self._h, C.byref(v)))
        return int(v.value)

    @property
    def sort_interval(self) -> int:
        return getattr(self, "_sort_interval", 100)

    @sort_interval.setter
    def sort_interval(self, v: int) -> None:
        check(load().rh_engine2_set_sort_interval(self._h, int(v)))
        self._sort_interval = int(v)

    def sort_now(self) -> None:
        check(load().rh_engine2_sort_now(self._h))

    def update_system(self, sys: System) -> None:
        cs = sys.to_c()
        check(load().rh_engine2_update_system(self._h, C.byref(cs)))
        self.sys = sys

    def kick_region(self, lo: Sequence[float], hi: Sequence[float], scale: float, add: Sequence[float]) -> None:
        lo_a, hi_a, add_a = (np.ascontiguousarray(x, np.float64) for x in (lo, hi, add))
        check(load().rh_engine2_kick_region(self._h, ptr(lo_a), ptr(hi_a), float(scale), ptr(add_a)))

    def set_species(self, ids: Sequence[int], species: int) -> None:
        a = np.ascontiguousarray(ids, np.uint32)
        check(load().rh_engine2_set_species(self._h, ptr(a), len(a), int(species)))

    def set_species_velocity(self, species: int, v: Sequence[float]) -> None:
        a = np.ascontiguousarray(v, np.float64)
        check(load().rh_engine2_set_species_velocity(self._h, int(species), ptr(a)))

    def read_positions(self) -> np.ndarray:
        pos = np.zeros((self._n, 2), np.int64)
        check(load().rh_engine2_read_positions(self._h, ptr(pos)))
        return pos

    def read_potential(self) -> float:
        v = C.c_double()
        check(load().rh_engine2_read_potential(self._h, C.byref(v)))
        return float(v.value)

    def read_ledger(self) -> Ledger:
        l = _lib.CLedger2()
        check(load().rh_engine2_read_ledger(self._h, C.byref(l)))
        return Ledger(np.array(l.energy[:]), np.array(l.impulse[:]), np.array(l.hits[:], dtype=np.uint64))

    def read_compartment_ledger(self, per_wall: int) -> CompartmentLedger:
        e = np.zeros((7, per_wall), np.float64)
        p = np.zeros((7, per_wall), np.float64)
        h = np.zeros((7, per_wall), np.uint64)
        check(load().rh_engine2_read_compartment_ledger(self._h, int(per_wall), ptr(e), ptr(p), ptr(h)))
        return CompartmentLedger(e, p, h)

    def read_piston(self) -> Tuple[float, float, float, float]:
        """(height sigma, velocity sigma/tau, gas force on it eps/sigma, work done on it by the gas eps)."""
        out = np.zeros(4, np.float64)
        check(load().rh_engine2_read_piston(self._h, ptr(out)))
        return tuple(float(x) for x in out)

    def set_piston_speed(self, u: float) -> None:
        check(load().rh_engine2_set_piston_speed(self._h, float(u)))

    def flip_bit(self, particle: int, axis: int, bit: int) -> None:
        check(load().rh_engine2_flip_bit(self._h, int(particle), int(axis), int(bit)))

    def flip_species_bit(self, species: int, axis: int, bit: int) -> None:
        check(load().rh_engine2_flip_species_bit(self._h, int(species), int(axis), int(bit)))

    def add_bond(self, a: int, b: int, ty: int) -> bool:
        c = C.c_int()
        check(load().rh_engine2_add_bond(self._h, int(a), int(b), int(ty), C.byref(c)))
        return bool(c.value)

    def remove_bond(self, a: int, b: int) -> bool:
        c = C.c_int()
        check(load().rh_engine2_remove_bond(self._h, int(a), int(b), C.byref(c)))
        return bool(c.value)
