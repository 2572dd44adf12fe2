"""md-core's data contract in Python: System3 / State3 (and the 2D System / State), Pair, Wall,
ForceMode and the seeded initial conditions.

Mirrors crates/md-core/src/{three,system,pair,init}.rs and md::ForceMode
(crates/md/src/engine.rs:104-142, engine3.rs:54-68) with the same names and argument meanings.
Initial conditions are computed by the engine library's host code (bit-identical to md-core).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check, load, ptr

ONE = 1 << 32  # one sigma in position units (fixed.rs:11)
UNIT = 1.0 / ONE
MAX_SPECIES = 4


def to_fixed(x: float) -> int:
    """fixed::to_fixed (fixed.rs:21-23): round half away from zero."""
    import math

    v = x * ONE
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def to_f64(x: int) -> float:
    return float(x) * UNIT


@dataclass(frozen=True)
class Pair:
    """md_core::Pair (pair.rs:3-8): truncated and shifted Lennard-Jones."""

    eps: float = 1.0
    sigma: float = 1.0
    rc: float = 2.5

    @staticmethod
    def lj(rc: float) -> "Pair":
        return Pair(1.0, 1.0, rc)

    @staticmethod
    def wca() -> "Pair":
        return Pair(1.0, 1.0, 2.0 ** (1.0 / 6.0))

    @staticmethod
    def none() -> "Pair":
        return Pair(0.0, 1.0, 0.0)


@dataclass(frozen=True)
class Wall:
    """md_core::Wall: kind 0 hard, 1 soft, 2 thermal (with temperature t)."""

    kind: int = 0
    t: float = 0.0

    HARD_KIND = 0
    SOFT_KIND = 1
    THERMAL_KIND = 2


Wall.Hard = Wall(0)  # type: ignore[attr-defined]
Wall.Soft = Wall(1)  # type: ignore[attr-defined]


def thermal_wall(t: float) -> Wall:
    return Wall(2, t)


@dataclass(frozen=True)
class Bath:
    t: float
    rate: float
    species: int = 0xFFFFFFFF


@dataclass
class System3:
    """md_core::three::System3 (three.rs:13-31)."""

    box_len: Tuple[int, int, int]
    dt: float
    masses: List[float]
    pairs: List[Pair]  # row-major [a * ns + b], symmetric
    walls: List[Optional[Tuple[Wall, Wall]]] = field(default_factory=lambda: [None, None, None])
    gravity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    seed: int = 1
    thermostat: Optional[Bath] = None
    pinned: int = 0

    @staticmethod
    def new(box_len, dt: float, pair: Pair, masses: Sequence[float]) -> "System3":
        if not (1 <= len(masses) <= MAX_SPECIES):
            raise ValueError("between 1 and 4 species")
        if pair.rc >= 7.9:
            raise ValueError("cutoff must be below 8 sigma")
        ns = len(masses)
        return System3(tuple(int(b) for b in box_len), float(dt), list(masses), [pair] * (ns * ns))

    def species(self) -> int:
        return len(self.masses)

    def with_pair(self, a: int, b: int, pair: Pair) -> "System3":
        if pair.rc >= 7.9:
            raise ValueError("cutoff must be below 8 sigma")
        ns = self.species()
        pairs = list(self.pairs)
        pairs[a * ns + b] = pair
        pairs[b * ns + a] = pair
        return replace(self, pairs=pairs)

    def with_walls(self, axis: int, lo: Wall, hi: Wall) -> "System3":
        walls = list(self.walls)
        walls[axis] = (lo, hi)
        return replace(self, walls=walls)

    def with_gravity(self, g) -> "System3":
        return replace(self, gravity=tuple(float(x) for x in g))

    def with_thermostat(self, t: float, rate: float) -> "System3":
        sp = self.thermostat.species if self.thermostat else 0xFFFFFFFF
        return replace(self, thermostat=Bath(t, rate, sp))

    def with_pinned(self, s: int) -> "System3":
        return replace(self, pinned=self.pinned | (1 << s))

    def with_seed(self, seed: int) -> "System3":
        return replace(self, seed=int(seed))

    def pair(self, a: int, b: int) -> Pair:
        return self.pairs[a * self.species() + b]

    def rc_max(self) -> float:
        return max(p.rc for p in self.pairs)

    def periodic(self, axis: int) -> bool:
        return self.walls[axis] is None

    def is_pinned(self, s: int) -> bool:
        return bool(self.pinned >> s & 1)

    def box_f64(self):
        return [to_f64(b) for b in self.box_len]

    def volume(self) -> float:
        b = self.box_f64()
        return b[0] * b[1] * b[2]

    def kick_coeff(self, s: int) -> float:
        return 0.0 if self.is_pinned(s) else self.dt * self.dt / self.masses[s] * ONE

    def velocity(self, v: int) -> float:
        return v * UNIT / self.dt

    def to_c(self) -> _lib.CSystem3:
        c = _lib.CSystem3()
        ns = self.species()
        for a in range(3):
            c.box_len[a] = int(self.box_len[a])
            w = self.walls[a]
            c.walled[a] = 0 if w is None else 1
            if w is not None:
                for side in range(2):
                    c.wall_kind[a][side] = w[side].kind
                    c.wall_t[a][side] = w[side].t
            c.gravity[a] = self.gravity[a]
        c.dt = self.dt
        c.nspecies = ns
        c.pinned = self.pinned
        for s in range(4):
            c.masses[s] = self.masses[s] if s < ns else 1.0
        for a in range(ns):
            for b in range(ns):
                p = self.pairs[a * ns + b]
                c.pairs[a * 4 + b] = _lib.Pair(p.eps, p.sigma, p.rc)
        c.seed = self.seed
        if self.thermostat is not None:
            c.has_bath = 1
            c.bath_species = self.thermostat.species & 0xFFFFFFFF
            c.bath_t = self.thermostat.t
            c.bath_rate = self.thermostat.rate
        return c


@dataclass
class State3:
    """md_core::three::State3 (three.rs:267-309): pos i64 (n,3), vel i32 (n,3), species u32 (n,)."""

    pos: np.ndarray
    vel: np.ndarray
    species: np.ndarray

    def __post_init__(self):
        self.pos = np.ascontiguousarray(self.pos, dtype=np.int64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(self.vel, dtype=np.int32).reshape(-1, 3)
        self.species = np.ascontiguousarray(self.species, dtype=np.uint32).reshape(-1)

    def n(self) -> int:
        return int(self.species.shape[0])

    def hash(self) -> int:
        """FNV-1a over every bit of positions, velocities and species (three.rs:280-297)."""
        return int(load().rh_hash3(self.n(), ptr(self.pos), ptr(self.vel), ptr(self.species)))

    def copy(self) -> "State3":
        return State3(self.pos.copy(), self.vel.copy(), self.species.copy())

    def __eq__(self, other) -> bool:
        return (
            isinstance(other, State3)
            and np.array_equal(self.pos, other.pos)
            and np.array_equal(self.vel, other.vel)
            and np.array_equal(self.species, other.species)
        )

    def half_step_kinetic(self, sys: System3) -> float:
        m = np.array([sys.masses[s] for s in range(sys.species())])[self.species]
        moving = np.array([not sys.is_pinned(int(s)) for s in self.species])
        v = self.vel.astype(np.float64) * UNIT / sys.dt
        return float(np.sum(0.5 * m[moving] * np.sum(v[moving] ** 2, axis=1)))


# ---- initial conditions (three.rs:311-404) ----

def cube_for(n: int, density: float) -> Tuple[int, int, int]:
    out = np.zeros(3, np.int64)
    check(load().rh_cube_for(n, density, ptr(out)))
    return tuple(int(x) for x in out)


def fcc_cells(n: int, box_len) -> Tuple[int, int, int]:
    b = np.ascontiguousarray(box_len, np.int64)
    out = np.zeros(3, np.uint64)
    check(load().rh_fcc_cells(n, ptr(b), ptr(out)))
    return tuple(int(x) for x in out)


def fcc(n: int, box_len, cells, jitter: float, seed: int) -> np.ndarray:
    b = np.ascontiguousarray(box_len, np.int64)
    c = np.ascontiguousarray(cells, np.uint64)
    pos = np.zeros((n, 3), np.int64)
    check(load().rh_fcc(n, ptr(b), ptr(c), jitter, seed, ptr(pos)))
    return pos


def maxwell(sys: System3, species, t: float, seed: int) -> np.ndarray:
    sp = np.ascontiguousarray(species, np.uint32)
    vel = np.zeros((len(sp), 3), np.int32)
    cs = sys.to_c()
    check(load().rh_maxwell3(C.byref(cs), len(sp), ptr(sp), t, seed, ptr(vel)))
    return vel


def set_temperature(state: State3, sys: System3, t: float) -> None:
    cs = sys.to_c()
    check(load().rh_set_temperature3(C.byref(cs), state.n(), ptr(state.species), ptr(state.vel), t))


def lj_liquid(n: int, density: float, t: float, seed_pos: int = 7, seed_vel: int = 3, jitter: float = 0.05,
              rc: float = 2.5, dt: float = 0.005) -> Tuple[System3, State3]:
    """An LJ liquid melted from an FCC lattice (as crates/md/tests/three.rs:15-31 builds it)."""
    sys = System3.new(cube_for(n, density), dt, Pair.lj(rc), [1.0])
    cells = fcc_cells(n, sys.box_len)
    pos = fcc(n, sys.box_len, cells, jitter, seed_pos)
    species = np.zeros(n, np.uint32)
    vel = maxwell(sys, species, t, seed_vel)
    return sys, State3(pos, vel, species)


# ---- 2D (system.rs, init.rs) ----

@dataclass(frozen=True)
class Thermostat:
    """md_core::Thermostat (system.rs:156-176): an Andersen bath at temperature t and collision
    rate `rate`, acting on the species in the bitmask `species` whose coordinate `axis` lies in
    [region[0], region[1]) (position units), redrawing around a mean x velocity `drift`."""

    t: float
    rate: float
    species: int = 0xFFFFFFFF
    region: Tuple[int, int] = (-(1 << 63), (1 << 63) - 1)
    axis: int = 0
    drift: float = 0.0


@dataclass(frozen=True)
class Membrane:
    """md_core::Membrane (system.rs:109-121): across `axis` at `position` (position units),
    blocking the species in the bitmask `blocks`; pores `pore_width` wide every 2^pore_shift."""

    axis: int
    position: int
    blocks: int
    pore_shift: int = 0
    pore_width: int = 0


@dataclass(frozen=True)
class Piston:
    """md_core::Piston (system.rs:95-107): a soft wall replacing the top wall.  speed is set for
    PistonDrive::Speed (sigma/tau); mass and force (pushing down, eps/sigma) for PistonDrive::Free."""

    start: float
    min: float
    max: float
    speed: float = 0.0
    free: bool = False
    mass: float = 0.0
    force: float = 0.0


@dataclass(frozen=True)
class BondType:
    """md_core::BondType (system.rs:70-93): Harmonic U = k/2 (r - r0)^2, or Fene
    U = -k R0^2/2 ln(1 - (r/R0)^2)."""

    fene: bool
    k: float
    r0: float

    @staticmethod
    def harmonic(k: float, r0: float) -> "BondType":
        return BondType(False, k, r0)

    @staticmethod
    def fene_bond(k: float, r0: float) -> "BondType":
        return BondType(True, k, r0)

    @staticmethod
    def kremer_grest() -> "BondType":
        return BondType(True, 30.0, 1.5)


@dataclass
class System:
    """md_core::System (system.rs:26-61) with its boundaries and builders (system.rs:186-353)."""

    box_len: Tuple[int, int]
    dt: float
    masses: List[float]
    pairs: List[Pair]
    walls: List[Optional[Tuple[Wall, Wall]]] = field(default_factory=lambda: [None, None])
    gravity: Tuple[float, float] = (0.0, 0.0)
    seed: int = 1
    pinned: int = 0
    compartments: Tuple[int, int] = (0, 0)
    diathermal: bool = False
    thermostat: Optional[Thermostat] = None
    membrane: Optional[Membrane] = None
    piston: Optional[Piston] = None
    bond_types: List[BondType] = field(default_factory=list)
    bonds: List[Tuple[int, int, int]] = field(default_factory=list)

    @staticmethod
    def new(box_len, dt: float, pair: Pair, masses: Sequence[float]) -> "System":
        ns = len(masses)
        if not (1 <= ns <= MAX_SPECIES):
            raise ValueError("between 1 and 4 species")
        return System(tuple(int(b) for b in box_len), float(dt), list(masses), [pair] * (ns * ns))

    def species(self) -> int:
        return len(self.masses)

    def with_pair(self, a: int, b: int, pair: Pair) -> "System":
        ns = self.species()
        pairs = list(self.pairs)
        pairs[a * ns + b] = pair
        pairs[b * ns + a] = pair
        return replace(self, pairs=pairs)

    def with_walls(self, axis: int, lo: Wall, hi: Wall) -> "System":
        walls = list(self.walls)
        walls[axis] = (lo, hi)
        return replace(self, walls=walls)

    def with_hard_walls(self, axis: int) -> "System":
        return self.with_walls(axis, Wall.Hard, Wall.Hard)

    @property
    def walled(self) -> Tuple[bool, bool]:
        return (self.walls[0] is not None, self.walls[1] is not None)

    def with_gravity(self, g) -> "System":
        return replace(self, gravity=tuple(float(x) for x in g))

    def with_pinned(self, s: int) -> "System":
        return replace(self, pinned=self.pinned | (1 << s))

    def with_seed(self, seed: int) -> "System":
        return replace(self, seed=int(seed))

    def with_compartments(self, axis: int, log2_sigma: int) -> "System":
        """Compartments of side 2^log2_sigma sigma along a walled axis (system.rs:233-239)."""
        shift = 32 + log2_sigma
        if self.walls[axis] is None:
            raise ValueError("compartments need walls on their axis")
        if self.box_len[axis] % (1 << shift):
            raise ValueError("box must be a whole number of compartments")
        c = list(self.compartments)
        c[axis] = shift
        return replace(self, compartments=tuple(c))

    def with_diathermal_compartments(self) -> "System":
        return replace(self, diathermal=True)

    def compartment_of(self, p) -> Tuple[int, int]:
        return tuple(0 if self.compartments[a] == 0 else int(p[a]) >> self.compartments[a] for a in range(2))

    def compartment_counts(self) -> Tuple[int, int]:
        return tuple(1 if self.compartments[a] == 0 else self.box_len[a] >> self.compartments[a] for a in range(2))

    def with_thermostat(self, t: float, rate: float) -> "System":
        th = self.thermostat
        if th is None:
            return replace(self, thermostat=Thermostat(t, rate))
        return replace(self, thermostat=replace(th, t=t, rate=rate))

    def with_fan(self, species: int, x0: float, x1: float, drift: float) -> "System":
        """The thermostat as a fan: species in [x0, x1) sigma, mean x velocity drift (system.rs:335-344)."""
        if self.thermostat is None:
            raise ValueError("a fan is a thermostat with a region: add the thermostat first")
        th = replace(self.thermostat, species=species, region=(to_fixed(x0), to_fixed(x1)), axis=0, drift=drift)
        return replace(self, thermostat=th)

    def with_bath_layer(self, y0: float, y1: float) -> "System":
        """The thermostat confined to y0 <= y < y1 sigma (system.rs:346-353)."""
        if self.thermostat is None:
            raise ValueError("a bath layer is a thermostat with a region: add the thermostat first")
        return replace(self, thermostat=replace(self.thermostat, region=(to_fixed(y0), to_fixed(y1)), axis=1))

    def with_membrane(self, axis: int, position_sigma: float, blocked: Sequence[int]) -> "System":
        if self.walls[axis] is None:
            raise ValueError("a membrane needs walls on its axis")
        blocks = 0
        for s in blocked:
            blocks |= 1 << s
        return replace(self, membrane=Membrane(axis, to_fixed(position_sigma), blocks))

    def with_pores(self, log2_period: int, width: float) -> "System":
        if self.membrane is None:
            raise ValueError("add a membrane first")
        return replace(self, membrane=replace(self.membrane, pore_shift=32 + log2_period, pore_width=to_fixed(width)))

    def with_piston(self, start: float, min_: float, speed: Optional[float] = None, mass: Optional[float] = None, force: float = 0.0) -> "System":
        """A piston at height start (sigma), stopping at min_ and the box height: driven at
        `speed` (sigma/tau), or free with `mass` under the external `force` (system.rs:284-291)."""
        if self.walls[1] is None:
            raise ValueError("a piston needs walls on y")
        top = to_f64(self.box_len[1])
        if not (start <= top and min_ < start):
            raise ValueError("piston needs min < start <= box height")
        if mass is not None:
            return replace(self, piston=Piston(start, min_, top, free=True, mass=mass, force=force))
        return replace(self, piston=Piston(start, min_, top, speed=float(speed or 0.0)))

    def with_bonds(self, types: Sequence[BondType], bonds: Sequence[Sequence[int]]) -> "System":
        if len(types) > 4:
            raise ValueError("at most 4 bond types")
        count = {}
        for b in bonds:
            if not (0 <= b[2] < len(types)):
                raise ValueError("bond type out of range")
            for i in (b[0], b[1]):
                count[i] = count.get(i, 0) + 1
                if count[i] > 4:
                    raise ValueError(f"particle {i} has more than 4 bonds")
        return replace(self, bond_types=list(types), bonds=[tuple(int(x) for x in b) for b in bonds])

    def rc_max(self) -> float:
        return max(p.rc for p in self.pairs)

    def periodic(self, axis: int) -> bool:
        return self.walls[axis] is None

    def is_pinned(self, s: int) -> bool:
        return bool(self.pinned >> s & 1)

    def box_f64(self):
        return [to_f64(b) for b in self.box_len]

    def to_step_displacement(self, v: float) -> int:
        x = round(v * self.dt * ONE)
        return max(-(1 << 31), min((1 << 31) - 1, int(x)))

    def to_c(self) -> _lib.CSystem2:
        """The C image; the bond array it points to is kept on the returned object."""
        c = _lib.CSystem2()
        ns = self.species()
        for a in range(2):
            c.box_len[a] = int(self.box_len[a])
            w = self.walls[a]
            c.walled[a] = 0 if w is None else 1
            if w is not None:
                for side in range(2):
                    c.wall_kind[a][side] = w[side].kind
                    c.wall_t[a][side] = w[side].t
            c.gravity[a] = self.gravity[a]
            c.compartments[a] = self.compartments[a]
        c.dt = self.dt
        c.nspecies = ns
        c.pinned = self.pinned
        for s in range(4):
            c.masses[s] = self.masses[s] if s < ns else 1.0
        for a in range(ns):
            for b in range(ns):
                p = self.pairs[a * ns + b]
                c.pairs[a * 4 + b] = _lib.Pair(p.eps, p.sigma, p.rc)
        c.seed = self.seed
        c.diathermal = 1 if self.diathermal else 0
        th = self.thermostat
        if th is not None:
            c.has_bath = 1
            c.bath_species = th.species & 0xFFFFFFFF
            c.bath_axis = th.axis
            c.bath_t = th.t
            c.bath_rate = th.rate
            c.bath_drift = th.drift
            c.bath_region[0] = th.region[0]
            c.bath_region[1] = th.region[1]
        m = self.membrane
        if m is not None:
            c.has_membrane = 1
            c.membrane_axis = m.axis
            c.membrane_blocks = m.blocks
            c.pore_shift = m.pore_shift
            c.membrane_position = m.position
            c.pore_width = m.pore_width
        pst = self.piston
        if pst is not None:
            c.has_piston = 1
            c.piston_free = 1 if pst.free else 0
            c.piston_speed = pst.speed
            c.piston_mass = pst.mass
            c.piston_force = pst.force
            c.piston_start = pst.start
            c.piston_min = pst.min
            c.piston_max = pst.max
        c.nbond_types = len(self.bond_types)
        for t, bt in enumerate(self.bond_types):
            c.bond_fene[t] = 1 if bt.fene else 0
            c.bond_k[t] = bt.k
            c.bond_r0[t] = bt.r0
        if self.bonds:
            arr = np.ascontiguousarray(np.array(self.bonds, dtype=np.uint32).reshape(-1, 3))
            c._bonds_keepalive = arr
            c.nbonds = arr.shape[0]
            c.bonds = arr.ctypes.data
        return c


@dataclass
class State:
    """md_core::State (system.rs:310-408): pos i64 (n,2), vel i32 (n,2), species u32 (n,)."""

    pos: np.ndarray
    vel: np.ndarray
    species: np.ndarray

    def __post_init__(self):
        self.pos = np.ascontiguousarray(self.pos, dtype=np.int64).reshape(-1, 2)
        self.vel = np.ascontiguousarray(self.vel, dtype=np.int32).reshape(-1, 2)
        self.species = np.ascontiguousarray(self.species, dtype=np.uint32).reshape(-1)

    def n(self) -> int:
        return int(self.species.shape[0])

    def hash(self) -> int:
        return int(load().rh_hash2(self.n(), ptr(self.pos), ptr(self.vel), ptr(self.species)))

    def copy(self) -> "State":
        return State(self.pos.copy(), self.vel.copy(), self.species.copy())

    def __eq__(self, other) -> bool:
        return (
            isinstance(other, State)
            and np.array_equal(self.pos, other.pos)
            and np.array_equal(self.vel, other.vel)
            and np.array_equal(self.species, other.species)
        )


def box_for(n: int, density: float, aspect: float = 1.0) -> Tuple[int, int]:
    out = np.zeros(2, np.int64)
    check(load().rh_box_for(n, density, aspect, ptr(out)))
    return int(out[0]), int(out[1])


def fluid(sys: System, n: int, t: float, seed: int) -> State:
    """init::fluid (init.rs:107-121): a jittered lattice with Maxwell velocities."""
    pos = np.zeros((n, 2), np.int64)
    vel = np.zeros((n, 2), np.int32)
    sp = np.zeros(n, np.uint32)
    cs = sys.to_c()
    check(load().rh_fluid2(C.byref(cs), n, t, seed, ptr(pos), ptr(vel), ptr(sp)))
    return State(pos, vel, sp)


def maxwell2(sys: System, species, t: float, seed: int) -> np.ndarray:
    sp = np.ascontiguousarray(species, np.uint32)
    vel = np.zeros((len(sp), 2), np.int32)
    cs = sys.to_c()
    check(load().rh_maxwell2(C.byref(cs), len(sp), ptr(sp), t, seed, ptr(vel)))
    return vel


# ---- ForceMode (engine.rs:104-142) ----

@dataclass(frozen=True)
class ForceMode:
    kind: int  # 0 CellList, 1 Verlet
    skin: float = 0.0
    cap: int = 0

    @staticmethod
    def cell_list() -> "ForceMode":
        return ForceMode(0)

    def to_c(self) -> _lib.CMode:
        return _lib.CMode(self.kind, self.cap, self.skin)


ForceMode.CellList = ForceMode(0)  # type: ignore[attr-defined]


def verlet3_with_skin(sys: System3, n: int, skin: float) -> ForceMode:
    """engine3.rs:61-68 without the 128 MiB WebGPU binding clamp."""
    m = _lib.CMode()
    cs = sys.to_c()
    check(load().rh_verlet_mode3(C.byref(cs), n, skin, C.byref(m)))
    return ForceMode(1, m.skin, m.cap)


def verlet3(sys: System3, n: int) -> ForceMode:
    return verlet3_with_skin(sys, n, 0.4)


def verlet_with_skin(sys: System, n: int, skin: float) -> ForceMode:
    m = _lib.CMode()
    cs = sys.to_c()
    check(load().rh_verlet_mode2(C.byref(cs), n, skin, C.byref(m)))
    return ForceMode(1, m.skin, m.cap)


def verlet_for(sys: System, n: int) -> ForceMode:
    return verlet_with_skin(sys, n, 0.4)


VERLET_MIN_N = 24_000  # ForceMode::VERLET_MIN_N (crates/md/src/engine.rs:116)


def force_mode_auto(sys: System, n: int) -> ForceMode:
    """ForceMode::auto (engine.rs:119-121): a cell list every step below VERLET_MIN_N
    particles, Verlet lists with a 0.4 sigma skin from there."""
    return ForceMode.cell_list() if n < VERLET_MIN_N else verlet_for(sys, n)


ForceMode.VERLET_MIN_N = VERLET_MIN_N  # type: ignore[attr-defined]
ForceMode.auto = staticmethod(force_mode_auto)  # type: ignore[attr-defined]
