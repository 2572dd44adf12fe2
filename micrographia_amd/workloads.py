"""Synthetic workloads: the batched ensemble and the serving request stream.

Each system's parameters are drawn from Threefry-2x32 counter streams (gpu-prims threefry.rs)
keyed by a meta-seed, so an ensemble is reproducible system by system.  Initial states use
md-core's initial conditions: 3D systems are FCC lattices with Maxwell velocities
(crates/md-core/src/three.rs:311-373), 2D systems jittered square lattices (init::fluid,
crates/md-core/src/init.rs:107-121).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple, Union

import numpy as np

from .core import Pair, State, State3, System, System3, box_for, cube_for, fcc, fcc_cells, fluid, maxwell

_ROT = (13, 15, 26, 6, 17, 29, 16, 24)
_M32 = 0xFFFFFFFF


def threefry2x32(c0: int, c1: int, k0: int, k1: int) -> Tuple[int, int]:
    """Threefry-2x32-20 (Random123; crates/gpu-prims/src/threefry.rs:7-22)."""
    ks = (k0 & _M32, k1 & _M32, (0x1BD11BDA ^ k0 ^ k1) & _M32)
    x0 = (c0 + ks[0]) & _M32
    x1 = (c1 + ks[1]) & _M32
    for r in range(20):
        x0 = (x0 + x1) & _M32
        rot = _ROT[r % 8]
        x1 = ((x1 << rot) | (x1 >> (32 - rot))) & _M32
        x1 ^= x0
        if r & 3 == 3:
            s = (r >> 2) + 1
            x0 = (x0 + ks[s % 3]) & _M32
            x1 = (x1 + ks[(s + 1) % 3] + s) & _M32
    return x0, x1


def uniform(seed: int, stream: int, i: int) -> float:
    """threefry::uniform: 53 random bits in [0, 1)."""
    r0, r1 = threefry2x32(i, stream, seed & _M32, seed >> 32)
    bits = ((r0 << 21) ^ (r1 >> 11)) & ((1 << 53) - 1)
    return bits / float(1 << 53)


@dataclass
class SystemSpec:
    index: int
    dim: int
    n: int
    density: float
    temperature: float
    seed: int
    steps: int = 0

    def build(self) -> Tuple[Union[System, System3], Union[State, State3]]:
        return make_system(self.dim, self.n, self.density, self.temperature, self.seed)


def make_system(dim: int, n: int, density: float, t: float, seed: int, rc: float = 2.5, dt: float = 0.005):
    """An LJ fluid of n particles: FCC + Maxwell in 3D, a jittered lattice in 2D."""
    if dim == 3:
        sys = System3.new(cube_for(n, density), dt, Pair.lj(rc), [1.0])
        pos = fcc(n, sys.box_len, fcc_cells(n, sys.box_len), 0.05, seed)
        sp = np.zeros(n, np.uint32)
        return sys, State3(pos, maxwell(sys, sp, t, seed), sp)
    sys = System.new(box_for(n, density, 1.0), dt, Pair.lj(rc), [1.0])
    return sys, fluid(sys, n, t, seed)


# streams of the ensemble draws
_S_DIM, _S_N, _S_RHO, _S_T, _S_SEED, _S_STEPS, _S_ARRIVAL = 11, 12, 13, 14, 15, 16, 17


def ensemble(count: int = 512, meta_seed: int = 2026, n_lo: int = 256, n_hi: int = 65536,
             rho=(0.3, 0.9), temp=(0.5, 2.0)) -> List[SystemSpec]:
    """The batched ensemble: dimension 2D/3D 50/50, N log-uniform in [n_lo, n_hi], density and
    temperature uniform, each system with its own seed."""
    out = []
    for k in range(count):
        dim = 3 if uniform(meta_seed, _S_DIM, k) < 0.5 else 2
        n = int(round(math.exp(math.log(n_lo) + (math.log(n_hi) - math.log(n_lo)) * uniform(meta_seed, _S_N, k))))
        density = rho[0] + (rho[1] - rho[0]) * uniform(meta_seed, _S_RHO, k)
        t = temp[0] + (temp[1] - temp[0]) * uniform(meta_seed, _S_T, k)
        seed = threefry2x32(k, _S_SEED, meta_seed & _M32, meta_seed >> 32)[0] | 1
        out.append(SystemSpec(k, dim, n, density, t, seed))
    return out


@dataclass
class Request(SystemSpec):
    arrival: float = 0.0


def request_stream(rate: float, duration: float, meta_seed: int = 7, n_lo: int = 1000, n_hi: int = 1_000_000,
                   steps_lo: int = 100, steps_hi: int = 10_000, rho=(0.3, 0.9), temp=(0.5, 2.0)) -> List[Request]:
    """Poisson arrivals at `rate` per second over `duration` seconds of simulation requests:
    N log-uniform in [n_lo, n_hi], dimension 50/50, steps uniform in [steps_lo, steps_hi],
    density and temperature drawn as in the ensemble."""
    reqs: List[Request] = []
    t = 0.0
    k = 0
    while True:
        u = uniform(meta_seed, _S_ARRIVAL, k)
        t += -math.log(1.0 - u) / rate
        if t >= duration:
            break
        dim = 3 if uniform(meta_seed, _S_DIM, k) < 0.5 else 2
        n = int(round(math.exp(math.log(n_lo) + (math.log(n_hi) - math.log(n_lo)) * uniform(meta_seed, _S_N, k))))
        density = rho[0] + (rho[1] - rho[0]) * uniform(meta_seed, _S_RHO, k)
        temperature = temp[0] + (temp[1] - temp[0]) * uniform(meta_seed, _S_T, k)
        steps = steps_lo + int((steps_hi - steps_lo + 1) * uniform(meta_seed, _S_STEPS, k))
        seed = threefry2x32(k, _S_SEED, meta_seed & _M32, meta_seed >> 32)[0] | 1
        reqs.append(Request(k, dim, n, density, temperature, seed, steps, t))
        k += 1
    return reqs
