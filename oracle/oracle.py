"""ctypes loader for the C oracle (test infrastructure only).

The oracle restates the reference's CPU twin (md-ref) and the md-core pieces it uses; the
tests, the smoke check and the benchmark's CPU-baseline leg load it as the checker.  Nothing in
the product package imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "librahman_oracle.so")

GOLDEN_INT_HASH = 0x3AAEFF680A2F0278  # crates/md/src/checks.rs:167


class Pair(C.Structure):
    _fields_ = [("eps", C.c_double), ("sigma", C.c_double), ("rc", C.c_double)]


class System3(C.Structure):
    _fields_ = [
        ("box_len", C.c_int64 * 3),
        ("dt", C.c_double),
        ("nspecies", C.c_uint32),
        ("pinned", C.c_uint32),
        ("masses", C.c_double * 4),
        ("pairs", Pair * 16),
        ("walled", C.c_uint32 * 3),
        ("wall_kind", (C.c_uint32 * 2) * 3),
        ("wall_t", (C.c_double * 2) * 3),
        ("gravity", C.c_double * 3),
        ("seed", C.c_uint64),
        ("has_bath", C.c_uint32),
        ("bath_species", C.c_uint32),
        ("bath_t", C.c_double),
        ("bath_rate", C.c_double),
    ]


class Ledger3(C.Structure):
    _fields_ = [("energy", C.c_double * 7), ("impulse", C.c_double * 7), ("hits", C.c_uint64 * 7)]


class System2(C.Structure):
    _fields_ = [
        ("box_len", C.c_int64 * 2),
        ("dt", C.c_double),
        ("nspecies", C.c_uint32),
        ("pinned", C.c_uint32),
        ("masses", C.c_double * 4),
        ("pairs", Pair * 16),
        ("walled", C.c_uint32 * 2),
        ("gravity", C.c_double * 2),
        ("seed", C.c_uint64),
        ("wall_kind", (C.c_uint32 * 2) * 2),
        ("wall_t", (C.c_double * 2) * 2),
        ("compartments", C.c_uint32 * 2),
        ("diathermal", C.c_uint32),
        ("has_bath", C.c_uint32),
        ("bath_species", C.c_uint32),
        ("bath_axis", C.c_uint32),
        ("bath_t", C.c_double),
        ("bath_rate", C.c_double),
        ("bath_drift", C.c_double),
        ("bath_region", C.c_int64 * 2),
        ("has_membrane", C.c_uint32),
        ("membrane_axis", C.c_uint32),
        ("membrane_blocks", C.c_uint32),
        ("pore_shift", C.c_uint32),
        ("membrane_position", C.c_int64),
        ("pore_width", C.c_int64),
        ("has_piston", C.c_uint32),
        ("piston_free", C.c_uint32),
        ("piston_speed", C.c_double),
        ("piston_mass", C.c_double),
        ("piston_force", C.c_double),
        ("piston_start", C.c_double),
        ("piston_min", C.c_double),
        ("piston_max", C.c_double),
        ("nbond_types", C.c_uint32),
        ("bond_fene", C.c_uint32 * 4),
        ("bond_k", C.c_double * 4),
        ("bond_r0", C.c_double * 4),
        ("nbonds", C.c_uint64),
        ("bonds", C.c_void_p),
    ]


class State2(C.Structure):
    """IntEngine's run state (int_engine.rs:23-48): clock, piston, ledgers."""

    _fields_ = [
        ("clock", C.c_uint32),
        ("piston_w", C.c_int64),
        ("piston_v", C.c_int32),
        ("piston_work", C.c_double),
        ("ledger", Ledger3),
        ("comp_energy", (C.c_double * 256) * 7),
        ("comp_impulse", (C.c_double * 256) * 7),
        ("comp_hits", (C.c_uint64 * 256) * 7),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        L = _lib
        P = C.c_void_p
        u64, u32, i64, dbl = C.c_uint64, C.c_uint32, C.c_int64, C.c_double
        L.orc_threefry2x32_rounds.argtypes = [u32, u32, u32, u32, u32, P]
        L.orc_to_unit_f64.argtypes = [u32, u32]
        L.orc_to_unit_f64.restype = dbl
        L.orc_uniform.argtypes = [u64, u32, u32]
        L.orc_uniform.restype = dbl
        L.orc_normal.argtypes = [u64, u32, u32]
        L.orc_normal.restype = dbl
        for f in ("orc_log", "orc_cos", "orc_sin"):
            getattr(L, f).argtypes = [dbl]
            getattr(L, f).restype = dbl
        L.orc_to_fixed.argtypes = [dbl]
        L.orc_to_fixed.restype = i64
        L.orc_to_f64.argtypes = [i64]
        L.orc_to_f64.restype = dbl
        L.orc_min_image.argtypes = [i64, i64, i64]
        L.orc_min_image.restype = i64
        L.orc_wrap.argtypes = [i64, i64]
        L.orc_wrap.restype = i64
        L.orc_sep_f32.argtypes = [i64, P]
        L.orc_sep_f32.restype = C.c_int
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_set_mutation.restype = None
        for f in ("orc_pair_shift",):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = dbl
        for f in ("orc_pair_energy", "orc_pair_f_over_r"):
            getattr(L, f).argtypes = [P, dbl]
            getattr(L, f).restype = dbl
        L.orc3_cube_for.argtypes = [u64, dbl, P]
        L.orc3_fcc_cells.argtypes = [u64, P, P]
        L.orc3_fcc.argtypes = [u64, P, P, dbl, u64, P]
        L.orc3_maxwell.argtypes = [P, u64, P, dbl, u64, P]
        L.orc3_zero_velocity_sum.argtypes = [P, u64, P, P]
        L.orc3_set_temperature.argtypes = [P, u64, P, P, dbl]
        L.orc3_hash.argtypes = [u64, P, P, P]
        L.orc3_hash.restype = u64
        L.orc3_cells_for.argtypes = [P, dbl, P]
        L.orc3_kicks.argtypes = [C.c_int, P, u64, P, P, P, P]
        L.orc3_kicks_brute.argtypes = [C.c_int, P, u64, P, P, P]
        L.orc3_step.argtypes = [C.c_int, P, u64, P, P, P, u64, P, P]
        L.orc3_reverse.argtypes = [C.c_int, P, u64, P, P, P]
        L.orc3_energy.argtypes = [C.c_int, P, u64, P, P, P, P, P]
        L.orc2_box_for.argtypes = [u64, dbl, dbl, P]
        L.orc2_lattice.argtypes = [u64, P, dbl, u64, P]
        L.orc2_maxwell.argtypes = [P, u64, P, dbl, u64, P]
        L.orc2_fluid.argtypes = [P, u64, dbl, u64, P, P, P]
        L.orc2_hash.argtypes = [u64, P, P, P]
        L.orc2_hash.restype = u64
        L.orc2_cells_for.argtypes = [P, dbl, P]
        L.orc2_kicks.argtypes = [C.c_int, P, u64, P, P, P, P]
        L.orc2_kicks_brute.argtypes = [C.c_int, P, u64, P, P, P]
        L.orc2_step.argtypes = [C.c_int, P, u64, P, P, P, u64]
        L.orc2_state_init.argtypes = [P, P]
        L.orc2_state_init.restype = None
        L.orc2_run.argtypes = [C.c_int, P, P, u64, P, P, P, u64]
        L.orc2_reverse_run.argtypes = [C.c_int, P, P, u64, P, P, P]
        L.orc2_forces.argtypes = [C.c_int, P, P, u64, P, P, P, P, P, P]
        L.orc2_reverse.argtypes = [C.c_int, P, u64, P, P, P]
        L.orc2_energy.argtypes = [C.c_int, P, u64, P, P, P, P, P]
        L.orc2_int_table_len.argtypes = [P]
        L.orc2_int_table_len.restype = u64
        L.orc2_int_tables.argtypes = [P, P, P]
        L.orc2_int_kick.argtypes = [P, P, P, u64, u32, u32, i64, i64, P]
        L.orc2_int_kick.restype = C.c_int
        L.orc2_golden_setup.argtypes = [P, P, P, P]
        L.orc2_set_fast_path.argtypes = [C.c_int]
        L.orc2_set_fast_path.restype = None
        L.orc3_draws.argtypes = [u64, dbl, dbl, dbl, u32, u64, u32, P]
        L.orc3_draws.restype = None
    return _lib


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


# ---------------------------------------------------------------------------------------------
# Threefry and libm
# ---------------------------------------------------------------------------------------------

def threefry(ctr, key, rounds=20):
    out = np.zeros(2, np.uint32)
    lib().orc_threefry2x32_rounds(ctr[0], ctr[1], key[0], key[1], rounds, ptr(out))
    return int(out[0]), int(out[1])


def uniform(seed, stream, i):
    return lib().orc_uniform(seed, stream, i)


def normal(seed, stream, i):
    return lib().orc_normal(seed, stream, i)


# ---------------------------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------------------------

def pair(eps=1.0, sigma=1.0, rc=2.5) -> Pair:
    return Pair(eps, sigma, rc)


def system3(box_len, dt=0.005, masses=(1.0,), pair_=None, seed=1) -> System3:
    s = System3()
    for a in range(3):
        s.box_len[a] = int(box_len[a])
    s.dt = dt
    s.nspecies = len(masses)
    for k, m in enumerate(masses):
        s.masses[k] = m
    p = pair_ if pair_ is not None else Pair(1.0, 1.0, 2.5)
    for k in range(16):
        s.pairs[k] = Pair(p.eps, p.sigma, p.rc)
    s.seed = seed
    return s


def set_pair3(s, a, b, p: Pair):
    s.pairs[a * 4 + b] = Pair(p.eps, p.sigma, p.rc)
    s.pairs[b * 4 + a] = Pair(p.eps, p.sigma, p.rc)


def cube_for(n, density):
    out = np.zeros(3, np.int64)
    lib().orc3_cube_for(n, density, ptr(out))
    return out


def fcc_cells(n, box_len):
    b = np.ascontiguousarray(box_len, np.int64)
    out = np.zeros(3, np.uint64)
    lib().orc3_fcc_cells(n, ptr(b), ptr(out))
    return out


def fcc(n, box_len, cells, jitter, seed):
    b = np.ascontiguousarray(box_len, np.int64)
    c = np.ascontiguousarray(cells, np.uint64)
    pos = np.zeros((n, 3), np.int64)
    lib().orc3_fcc(n, ptr(b), ptr(c), jitter, seed, ptr(pos))
    return pos


def maxwell3(sys, species, t, seed):
    sp = np.ascontiguousarray(species, np.uint32)
    vel = np.zeros((len(sp), 3), np.int32)
    lib().orc3_maxwell(C.byref(sys), len(sp), ptr(sp), t, seed, ptr(vel))
    return vel


def set_temperature3(sys, vel, species, t):
    sp = np.ascontiguousarray(species, np.uint32)
    lib().orc3_set_temperature(C.byref(sys), len(sp), ptr(sp), ptr(vel), t)


def hash3(pos, vel, species) -> int:
    pos = np.ascontiguousarray(pos, np.int64)
    vel = np.ascontiguousarray(vel, np.int32)
    sp = np.ascontiguousarray(species, np.uint32)
    return int(lib().orc3_hash(len(sp), ptr(pos), ptr(vel), ptr(sp)))


def set_mutation(m: int) -> None:
    """Plant mistake m in the 3D pair loop (0 restores the true law); tests only."""
    lib().orc_set_mutation(int(m))


def cells_for3(sys, r):
    out = np.zeros(3, np.uint32)
    lib().orc3_cells_for(C.byref(sys), r, ptr(out))
    return out


def kicks3(sys, pos, species, prec=32):
    pos = np.ascontiguousarray(pos, np.int64)
    sp = np.ascontiguousarray(species, np.uint32)
    k = np.zeros((len(sp), 3), np.int32)
    pe = C.c_double(0.0)
    _check(lib().orc3_kicks(prec, C.byref(sys), len(sp), ptr(pos), ptr(sp), ptr(k), C.byref(pe)), "kicks3")
    return k, pe.value


def kicks3_brute(sys, pos, species, prec=32):
    pos = np.ascontiguousarray(pos, np.int64)
    sp = np.ascontiguousarray(species, np.uint32)
    k = np.zeros((len(sp), 3), np.int32)
    _check(lib().orc3_kicks_brute(prec, C.byref(sys), len(sp), ptr(pos), ptr(sp), ptr(k)), "kicks3_brute")
    return k


def step3(sys, pos, vel, species, steps, prec=32, clock=0):
    """In place; returns (clock, ledger)."""
    sp = np.ascontiguousarray(species, np.uint32)
    clk = C.c_uint32(clock)
    led = Ledger3()
    _check(lib().orc3_step(prec, C.byref(sys), len(sp), ptr(pos), ptr(vel), ptr(sp), steps, C.byref(clk), C.byref(led)), "step3")
    return clk.value, led


def reverse3(sys, pos, vel, species, prec=32):
    sp = np.ascontiguousarray(species, np.uint32)
    _check(lib().orc3_reverse(prec, C.byref(sys), len(sp), ptr(pos), ptr(vel), ptr(sp)), "reverse3")


def energy3(sys, pos, vel, species, prec=64):
    pos = np.ascontiguousarray(pos, np.int64)
    vel = np.ascontiguousarray(vel, np.int32)
    sp = np.ascontiguousarray(species, np.uint32)
    ke, pe = C.c_double(), C.c_double()
    _check(lib().orc3_energy(prec, C.byref(sys), len(sp), ptr(pos), ptr(vel), ptr(sp), C.byref(ke), C.byref(pe)), "energy3")
    return ke.value, pe.value


def liquid3(n, density, t, seed_pos=7, seed_vel=3, jitter=0.05, dt=0.005, rc=2.5):
    """An LJ liquid melted from FCC, as crates/md/tests/three.rs:15-31 builds it."""
    box = cube_for(n, density)
    sys = system3(box, dt=dt, pair_=Pair(1.0, 1.0, rc))
    cells = fcc_cells(n, box)
    pos = fcc(n, box, cells, jitter, seed_pos)
    species = np.zeros(n, np.uint32)
    vel = maxwell3(sys, species, t, seed_vel)
    return sys, pos, vel, species


# ---------------------------------------------------------------------------------------------
# 2D
# ---------------------------------------------------------------------------------------------

def system2(box_len, dt=0.005, masses=(1.0,), pair_=None, seed=1) -> System2:
    s = System2()
    for a in range(2):
        s.box_len[a] = int(box_len[a])
    s.dt = dt
    s.nspecies = len(masses)
    for k, m in enumerate(masses):
        s.masses[k] = m
    p = pair_ if pair_ is not None else Pair(1.0, 1.0, 2.5)
    for k in range(16):
        s.pairs[k] = Pair(p.eps, p.sigma, p.rc)
    s.seed = seed
    return s


def box_for(n, density, aspect=1.0):
    out = np.zeros(2, np.int64)
    lib().orc2_box_for(n, density, aspect, ptr(out))
    return out


def fluid2(sys, n, t, seed):
    pos = np.zeros((n, 2), np.int64)
    vel = np.zeros((n, 2), np.int32)
    sp = np.zeros(n, np.uint32)
    lib().orc2_fluid(C.byref(sys), n, t, seed, ptr(pos), ptr(vel), ptr(sp))
    return pos, vel, sp


def maxwell2(sys, species, t, seed):
    sp = np.ascontiguousarray(species, np.uint32)
    vel = np.zeros((len(sp), 2), np.int32)
    lib().orc2_maxwell(C.byref(sys), len(sp), ptr(sp), t, seed, ptr(vel))
    return vel


def hash2(pos, vel, species) -> int:
    pos = np.ascontiguousarray(pos, np.int64)
    vel = np.ascontiguousarray(vel, np.int32)
    sp = np.ascontiguousarray(species, np.uint32)
    return int(lib().orc2_hash(len(sp), ptr(pos), ptr(vel), ptr(sp)))


def kicks2(sys, pos, species, prec=32):
    pos = np.ascontiguousarray(pos, np.int64)
    sp = np.ascontiguousarray(species, np.uint32)
    k = np.zeros((len(sp), 2), np.int32)
    pe = C.c_double(0.0)
    _check(lib().orc2_kicks(prec, C.byref(sys), len(sp), ptr(pos), ptr(sp), ptr(k), C.byref(pe)), "kicks2")
    return k, pe.value


def kicks2_brute(sys, pos, species, prec=32):
    pos = np.ascontiguousarray(pos, np.int64)
    sp = np.ascontiguousarray(species, np.uint32)
    k = np.zeros((len(sp), 2), np.int32)
    _check(lib().orc2_kicks_brute(prec, C.byref(sys), len(sp), ptr(pos), ptr(sp), ptr(k)), "kicks2_brute")
    return k


def draws3(seed, t, mass, dt, id0, count, clock):
    """The stochastic boundaries' draws for keys (id0 + k, clock), as integer displacements:
    an array (count, 6) (bath x, y, z; heat-bath wall normal and two tangentials)."""
    out = np.zeros((count, 6), np.int32)
    lib().orc3_draws(seed, t, mass, dt, id0, count, clock, ptr(out))
    return out


class fast_path:
    """Within the block the oracle's 2D forces take md-ref's fast pair path (fast.rs), as
    IntEngine does by default, instead of the reference loop."""

    def __enter__(self):
        lib().orc2_set_fast_path(1)
        return self

    def __exit__(self, *exc):
        lib().orc2_set_fast_path(0)
        return False


def step2(sys, pos, vel, species, steps, prec=32):
    sp = np.ascontiguousarray(species, np.uint32)
    _check(lib().orc2_step(prec, C.byref(sys), len(sp), ptr(pos), ptr(vel), ptr(sp), steps), "step2")


def reverse2(sys, pos, vel, species, prec=32):
    sp = np.ascontiguousarray(species, np.uint32)
    _check(lib().orc2_reverse(prec, C.byref(sys), len(sp), ptr(pos), ptr(vel), ptr(sp)), "reverse2")


def energy2(sys, pos, vel, species, prec=64):
    pos = np.ascontiguousarray(pos, np.int64)
    vel = np.ascontiguousarray(vel, np.int32)
    sp = np.ascontiguousarray(species, np.uint32)
    ke, pe = C.c_double(), C.c_double()
    _check(lib().orc2_energy(prec, C.byref(sys), len(sp), ptr(pos), ptr(vel), ptr(sp), C.byref(ke), C.byref(pe)), "energy2")
    return ke.value, pe.value


def state2(sys) -> State2:
    st = State2()
    lib().orc2_state_init(C.byref(sys), C.byref(st))
    return st


def run2(sys, st: State2, pos, vel, species, steps, prec=32):
    """IntEngine::step with the full 2D boundaries; pos, vel in place, st advanced."""
    sp = np.ascontiguousarray(species, np.uint32)
    _check(lib().orc2_run(prec, C.byref(sys), C.byref(st), len(sp), ptr(pos), ptr(vel), ptr(sp), steps), "run2")


def reverse_run2(sys, st: State2, pos, vel, species, prec=32):
    sp = np.ascontiguousarray(species, np.uint32)
    _check(lib().orc2_reverse_run(prec, C.byref(sys), C.byref(st), len(sp), ptr(pos), ptr(vel), ptr(sp)), "reverse_run2")


def forces2(sys, st: State2, pos, species, prec=32):
    """(kicks, pe, wall_force (n, 7), piston_force)"""
    pos = np.ascontiguousarray(pos, np.int64)
    sp = np.ascontiguousarray(species, np.uint32)
    n = len(sp)
    k = np.zeros((n, 2), np.int32)
    wf = np.zeros((n, 7), np.float64)
    pe, pf = C.c_double(), C.c_double()
    _check(lib().orc2_forces(prec, C.byref(sys), C.byref(st), n, ptr(pos), ptr(sp), ptr(k), C.byref(pe), ptr(wf), C.byref(pf)), "forces2")
    return k, pe.value, wf, pf.value


def golden_setup():
    sys = System2()
    pos = np.zeros((3000, 2), np.int64)
    vel = np.zeros((3000, 2), np.int32)
    sp = np.zeros(3000, np.uint32)
    lib().orc2_golden_setup(C.byref(sys), ptr(pos), ptr(vel), ptr(sp))
    return sys, pos, vel, sp


def int_tables(sys):
    ln = int(lib().orc2_int_table_len(C.byref(sys)))
    ns = sys.nspecies
    data = np.zeros(ns * ns * ln, np.int32)
    rc2 = np.zeros(ns * ns, np.uint32)
    lib().orc2_int_tables(C.byref(sys), ptr(data), ptr(rc2))
    return data, rc2, ln


def int_kick(sys, tables, sa, sb, dx, dy):
    data, rc2, ln = tables
    out = np.zeros(2, np.int32)
    ok = lib().orc2_int_kick(C.byref(sys), ptr(data), ptr(rc2), ln, sa, sb, int(dx), int(dy), ptr(out))
    return (int(out[0]), int(out[1])) if ok else None
