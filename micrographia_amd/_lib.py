"""ctypes binding of librahman_hip.so (the C ABI in include/rahman_hip.h).

The shared library is built in-tree by ``micrographia_amd.build``; importing this module loads it
and fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librahman_hip.so")

RH_OK = 0
RH_ERR_INVALID = -1      # a contract violation (the reference's assert!)
RH_ERR_HIP = -2          # a HIP runtime failure
RH_ERR_NOMEM = -3        # allocation failed or over budget
RH_ERR_UNSUPPORTED = -4  # a system or command this engine does not run


class RahmanError(RuntimeError):
    """A failed call into the engine (the reference's assert!/panic, as a Python exception)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class Pair(C.Structure):
    _fields_ = [("eps", C.c_double), ("sigma", C.c_double), ("rc", C.c_double)]


class CSystem3(C.Structure):
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


class CSystem2(C.Structure):
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


class CMode(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("cap", C.c_uint32), ("skin", C.c_double)]


class CLedger3(C.Structure):
    _fields_ = [("energy", C.c_double * 7), ("impulse", C.c_double * 7), ("hits", C.c_uint64 * 7)]


class CLedger2(C.Structure):
    _fields_ = [("energy", C.c_double * 7), ("impulse", C.c_double * 7), ("hits", C.c_uint64 * 7)]


class CListStats(C.Structure):
    _fields_ = [
        ("rebuilds", C.c_uint32),
        ("overflowed", C.c_uint32),
        ("cap", C.c_uint32),
        ("max_count", C.c_uint32),
        ("steps_done", C.c_uint64),
        ("launches_skipped", C.c_uint64),
    ]


# Every function the header declares, with its argument and result types.
P = C.c_void_p
u32, u64, i64, dbl, i32 = C.c_uint32, C.c_uint64, C.c_int64, C.c_double, C.c_int
SIGNATURES = {
    "rh_last_error": ([], C.c_char_p),
    "rh_version": ([], C.c_char_p),
    "rh_device_count": ([P], i32),
    "rh_cube_for": ([u64, dbl, P], i32),
    "rh_fcc_cells": ([u64, P, P], i32),
    "rh_fcc": ([u64, P, P, dbl, u64, P], i32),
    "rh_maxwell3": ([P, u64, P, dbl, u64, P], i32),
    "rh_set_temperature3": ([P, u64, P, P, dbl], i32),
    "rh_hash3": ([u64, P, P, P], u64),
    "rh_box_for": ([u64, dbl, dbl, P], i32),
    "rh_lattice": ([u64, P, dbl, u64, P], i32),
    "rh_maxwell2": ([P, u64, P, dbl, u64, P], i32),
    "rh_fluid2": ([P, u64, dbl, u64, P, P, P], i32),
    "rh_hash2": ([u64, P, P, P], u64),
    "rh_verlet_mode3": ([P, u64, dbl, P], i32),
    "rh_verlet_mode2": ([P, u64, dbl, P], i32),
    "rh_engine3_create": ([P, u64, P, P, P, P, P], i32),
    "rh_engine3_bytes_needed": ([P, u64, P, P], i32),
    "rh_engine3_device_bytes": ([P, P], i32),
    "rh_engine3_destroy": ([P], i32),
    "rh_engine3_n": ([P, P], i32),
    "rh_engine3_steps": ([P, P], i32),
    "rh_engine3_step": ([P, u64], i32),
    "rh_engine3_reverse": ([P], i32),
    "rh_engine3_update_system": ([P, P], i32),
    "rh_engine3_kick_region": ([P, P, P, dbl, P], i32),
    "rh_engine3_set_species": ([P, P, u64, u32], i32),
    "rh_engine3_set_species_velocity": ([P, u32, P], i32),
    "rh_engine3_upload": ([P, P, P, P], i32),
    "rh_engine3_read_state": ([P, P, P, P], i32),
    "rh_engine3_kicks": ([P, P, P], i32),
    "rh_engine3_measure_kinetic": ([P], i32),
    "rh_engine3_read_kinetic": ([P, P, P], i32),
    "rh_engine3_read_kinetic_total": ([P, P], i32),
    "rh_engine3_read_potential": ([P, P], i32),
    "rh_engine3_read_ledger": ([P, P], i32),
    "rh_engine3_list_stats": ([P, P], i32),
    "rh_engine3_set_sort_interval": ([P, u32], i32),
    "rh_engine3_tune": ([P, i32, u32], i32),
    "rh_engine3_stream": ([P, P], i32),
    "rh_engine3_set_timing": ([P, i32], i32),
    "rh_engine3_kernel_ms": ([P, P, P, P, P], i32),
    "rh_engine2_create": ([P, u64, P, P, P, P, P], i32),
    "rh_engine2_bytes_needed": ([P, u64, P, P], i32),
    "rh_engine2_device_bytes": ([P, P], i32),
    "rh_engine2_destroy": ([P], i32),
    "rh_engine2_step": ([P, u64], i32),
    "rh_engine2_reverse": ([P], i32),
    "rh_engine2_upload": ([P, P, P, P], i32),
    "rh_engine2_read_state": ([P, P, P, P], i32),
    "rh_engine2_compute_forces": ([P], i32),
    "rh_engine2_kicks": ([P, P, P], i32),
    "rh_engine2_read_kinetic_total": ([P, P], i32),
    "rh_engine2_list_stats": ([P, P], i32),
    "rh_engine2_set_arith": ([P, i32], i32),
    "rh_engine2_tune": ([P, i32, u32], i32),
    "rh_engine2_stream": ([P, P], i32),
    "rh_engine2_n": ([P, P], i32),
    "rh_engine2_steps": ([P, P], i32),
    "rh_engine2_update_system": ([P, P], i32),
    "rh_engine2_set_sort_interval": ([P, u32], i32),
    "rh_engine2_sort_now": ([P], i32),
    "rh_engine2_read_potential": ([P, P], i32),
    "rh_engine2_read_positions": ([P, P], i32),
    "rh_engine2_kick_region": ([P, P, P, dbl, P], i32),
    "rh_engine2_set_species": ([P, P, u64, u32], i32),
    "rh_engine2_set_species_velocity": ([P, u32, P], i32),
    "rh_engine2_read_ledger": ([P, P], i32),
    "rh_engine2_read_compartment_ledger": ([P, u32, P, P, P], i32),
    "rh_engine2_read_piston": ([P, P], i32),
    "rh_engine2_set_piston_speed": ([P, dbl], i32),
    "rh_engine2_flip_bit": ([P, u32, u32, u32], i32),
    "rh_engine2_flip_species_bit": ([P, u32, u32, u32], i32),
    "rh_engine2_add_bond": ([P, u32, u32, u32, P], i32),
    "rh_engine2_remove_bond": ([P, u32, u32, P], i32),
    "rh_batch_create": ([u64, u32, P], i32),
    "rh_batch_destroy": ([P], i32),
    "rh_batch_add3": ([P, P, u64, P, P, P, dbl, u64, P], i32),
    "rh_batch_add2": ([P, P, u64, P, P, P, dbl, u64, P], i32),
    "rh_batch_remove": ([P, C.c_int32], i32),
    "rh_batch_set_steps": ([P, C.c_int32, u64], i32),
    "rh_batch_run": ([P, u64], i32),
    "rh_batch_sync": ([P], i32),
    "rh_batch_poll": ([P, P, u32], i32),
    "rh_batch_read_state": ([P, C.c_int32, P, P, P], i32),
    "rh_batch_list_stats": ([P, C.c_int32, P], i32),
    "rh_batch_read_ledger": ([P, C.c_int32, P], i32),
    "rh_batch_set_graph_chunk": ([P, u32], i32),
    "rh_batch_set_tuning": ([P, u32, u32, u32, dbl, u32], i32),
    "rh_batch_stream": ([P, P], i32),
    "rh_batch_capacity": ([P, P, P, P, P], i32),
    "rh_batch_bytes_needed3": ([P, P, u64, dbl, P], i32),
    "rh_batch_bytes_needed2": ([P, P, u64, dbl, P], i32),
    "rh_prims_exclusive_scan_u32": ([P, P, u64, P], i32),
    "rh_prims_radix_sort_u32": ([P, P, u64, u32, P], i32),
    "rh_prims_radix_sort_u32_noalloc": ([P, P, u64, u32, P], i32),
    "rh_prims_reduce_sum_f64": ([P, P, u64, P], i32),
    "rh_prims_threefry": ([P, u64, u32, u32, u32, u32, P], i32),
    "rh_prims_atomic_add_i64": ([P, u64, P, P], i32),
    "rh_prims_stochastic_draws": ([u64, dbl, dbl, dbl, u32, u64, u32, P, P], i32),
    "rh_prims_kick_convert": ([P, P, u64, P], i32),
    "rh_pair_division_floor": ([dbl, dbl, dbl, P], i32),
    "rh_prims_check_pair_division": ([dbl, dbl, dbl, P, P, P, P], i32),
}

_lib = None


def load() -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m micrographia_amd.build` "
                "(the engine has no CPU fallback)"
            )
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (same SONAME), and a
        # process that loads this library first and torch later ends up with two runtimes, the
        # second of which sees no device.  Loading torch first makes this library bind to the
        # runtime torch already mapped.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def check(code: int) -> None:
    if code != RH_OK:
        msg = load().rh_last_error()
        raise RahmanError(code, msg.decode() if msg else "unknown error")


def ptr(a):
    """ctypes pointer to a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the engine must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)
