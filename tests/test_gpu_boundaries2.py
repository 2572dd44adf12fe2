"""The 2D engine's boundaries on the GPU against the oracle's IntEngine<f32> with md_core::System's
full boundaries (crates/md-ref/src/int_engine.rs:130-447): compartments, membrane and pores,
thermostat regions (fan, bath layer), heat-bath and soft walls, the piston (driven and free),
bonds, the ledger per wall and compartment, and the run-time commands (bit flips, bonds added and
removed, update_system).  Trajectories must be identical bit for bit in both force modes; ledger
sums agree to their fixed-point rounding (2^-32 per booking)."""
import ctypes as C

import numpy as np
import pytest

import micrographia_amd as rh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ONE = 1 << 32
MODES = ["cells", "verlet"]


def oracle_system(sys: rh.System) -> O.System2:
    """The same system in the oracle's struct (field by field; the bond array stays alive on it)."""
    c = sys.to_c()
    o = O.System2()
    for name, _ in O.System2._fields_:
        src, dst = getattr(type(c), name), getattr(O.System2, name)
        assert src.size == dst.size, name
        C.memmove(C.addressof(o) + dst.offset, C.addressof(c) + src.offset, src.size)
    o._keep = c
    return o


def engine(sys, st, mode):
    m = rh.ForceMode.cell_list() if mode == "cells" else rh.verlet_for(sys, st.n())
    return rh.GpuEngine(sys, st, m)


class Twin:
    """The oracle's run of the same system and state."""

    def __init__(self, sys, st):
        self.sys = oracle_system(sys)
        self.st = O.state2(self.sys)
        self.pos, self.vel, self.sp = st.pos.copy(), st.vel.copy(), st.species.copy()

    def run(self, steps):
        O.run2(self.sys, self.st, self.pos, self.vel, self.sp, steps)

    def reverse(self):
        O.reverse_run2(self.sys, self.st, self.pos, self.vel, self.sp)

    def state(self):
        return rh.State(self.pos, self.vel, self.sp)

    def update(self, sys):
        self.sys = oracle_system(sys)


def walled_fluid(lx, ly, n, t=1.0, seed=3, masses=(1.0,), walls_x=True, fill_y=None):
    """A fluid in an lx x ly sigma box with hard walls on y (and on x unless walls_x is False),
    filling the lowest fill_y sigma (all of it by default)."""
    sys = rh.System.new((lx * ONE, ly * ONE), 0.005, rh.Pair.lj(2.5), list(masses)).with_hard_walls(1)
    if walls_x:
        sys = sys.with_hard_walls(0)
    if fill_y is None:
        return sys, rh.fluid(sys, n, t, seed)
    low = rh.System.new((lx * ONE, fill_y * ONE), 0.005, rh.Pair.lj(2.5), list(masses)).with_hard_walls(1)
    return sys, rh.fluid(low, n, t, seed)


def close(a, b, tol=1e-6):
    return abs(a - b) <= tol + 1e-9 * max(abs(a), abs(b))


def check_ledger(e, twin, per_wall=1):
    """Totals and the per-compartment ledger against the oracle's."""
    led = e.read_ledger()
    ol = twin.st.ledger
    for w in range(7):
        assert int(led.hits[w]) == int(ol.hits[w]), w
        assert close(led.energy[w], ol.energy[w], 1e-6 + 1e-9 * int(ol.hits[w] + 1)), (w, led.energy[w], ol.energy[w])
        assert close(led.impulse[w], ol.impulse[w], 1e-6 + 1e-9 * int(ol.hits[w] + 1)), (w, led.impulse[w], ol.impulse[w])
    cl = e.read_compartment_ledger(per_wall)
    for w in range(7):
        for c in range(per_wall):
            assert int(cl.hits[w, c]) == int(twin.st.comp_hits[w][c]), (w, c)
            assert close(cl.impulse[w, c], twin.st.comp_impulse[w][c], 1e-5), (w, c)
            assert close(cl.energy[w, c], twin.st.comp_energy[w][c], 1e-5), (w, c)
    return led


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("diathermal", [False, True])
def test_compartments(mode, diathermal):
    sys, st = walled_fluid(64, 32, 900, t=1.5)
    sys = sys.with_compartments(0, 4).with_walls(1, rh.Wall.Soft, rh.Wall.Hard)  # four 16-sigma boxes, a soft floor
    if diathermal:
        sys = sys.with_diathermal_compartments()
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(150)
    e.step(150)
    twin.run(300)
    assert e.read_state() == twin.state()
    led = check_ledger(e, twin, per_wall=4)
    assert led.hits[2] == 0 and led.impulse[2] > 0  # the soft floor pushes without hits
    home = st.pos[:, 0] >> 36
    assert np.array_equal(e.read_state().pos[:, 0] >> 36, home)


@pytest.mark.parametrize("mode", MODES)
def test_membrane_and_pores(mode):
    sys, st = walled_fluid(48, 48, 1200, t=1.5, masses=(1.0, 2.0))
    sp = st.species.copy()
    sp[::2] = 1
    st = rh.State(st.pos, rh.maxwell2(sys, sp, 1.5, 4), sp)
    keep = ~((sp == 1) & (st.pos[:, 0] >= 24 * ONE))
    st = rh.State(st.pos[keep], st.vel[keep], st.species[keep])
    sys = sys.with_membrane(0, 24.0, [1])
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(300)
    twin.run(300)
    assert e.read_state() == twin.state()
    led = check_ledger(e, twin)
    assert led.hits[4] + led.hits[5] > 0
    # pores every 8 sigma, 3 sigma wide
    sys2 = sys.with_pores(3, 3.0)
    e2, twin2 = engine(sys2, e.read_state(), mode), Twin(sys2, e.read_state())
    e2.step(400)
    twin2.run(400)
    s2 = e2.read_state()
    assert s2 == twin2.state()
    assert (s2.pos[s2.species == 1, 0] >= 24 * ONE).any()


@pytest.mark.parametrize("mode", MODES)
def test_fan_and_bath_layer(mode):
    sys, st = walled_fluid(60, 30, 800, t=0.5, walls_x=False)
    fan = sys.with_thermostat(0.5, 5.0).with_fan(0xFFFFFFFF, 0.0, 20.0, 1.0)
    e, twin = engine(fan, st, mode), Twin(fan, st)
    e.step(400)
    twin.run(400)
    s = e.read_state()
    assert s == twin.state()
    led = check_ledger(e, twin)
    assert led.hits[6] > 100
    assert (s.vel[:, 0] / (0.005 * ONE)).mean() > 0.05
    layer = sys.with_thermostat(2.0, 10.0).with_bath_layer(0.0, 5.0)
    e, twin = engine(layer, st, mode), Twin(layer, st)
    e.step(300)
    twin.run(300)
    assert e.read_state() == twin.state()
    check_ledger(e, twin)


@pytest.mark.parametrize("mode", MODES)
def test_heat_bath_and_soft_walls(mode):
    sys, st = walled_fluid(40, 40, 600, t=1.0)
    sys = sys.with_walls(1, rh.thermal_wall(2.0), rh.thermal_wall(1.0)).with_walls(0, rh.Wall.Soft, rh.Wall.Soft).with_gravity((0.0, -0.1))
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(800)
    twin.run(800)
    assert e.read_state() == twin.state()
    led = check_ledger(e, twin)
    assert led.hits[2] + led.hits[3] > 10 and led.impulse[0] > 0 and led.impulse[1] > 0


def check_piston(e, twin):
    h, v, f, w = e.read_piston()
    assert h == twin.st.piston_w / ONE
    assert v == twin.st.piston_v / ONE / 0.005
    assert close(w, twin.st.piston_work, 1e-9)
    return h, v, f, w


@pytest.mark.parametrize("mode", MODES)
def test_driven_piston_with_thermostat(mode):
    sys, st = walled_fluid(40, 40, 500, t=1.0, fill_y=36)
    sys = sys.with_piston(38.0, 25.0, speed=-0.5).with_thermostat(1.0, 2.0)
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(200)
    twin.run(200)
    assert e.read_state() == twin.state()
    check_piston(e, twin)
    e.set_piston_speed(0.25)
    twin.st.piston_v = sys.to_step_displacement(0.25)
    e.step(200)
    twin.run(200)
    assert e.read_state() == twin.state()
    h, v, _, _ = check_piston(e, twin)
    assert v == pytest.approx(0.25, rel=1e-6) or h == 40.0
    check_ledger(e, twin)


@pytest.mark.parametrize("mode", MODES)
def test_free_piston_and_its_reversal(mode):
    sys, st = walled_fluid(30, 60, 300, fill_y=55)
    sys = sys.with_piston(58.0, 5.0, mass=50.0, force=20.0)
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(600)
    twin.run(600)
    assert e.read_state() == twin.state()
    h, v, f, w = check_piston(e, twin)
    assert h < 58.0 and w < 0.0
    # the piston is part of the reversible dynamics
    e.reverse()
    e.step(600)
    e.reverse()
    twin.reverse()
    twin.run(600)
    twin.reverse()
    assert e.read_state() == twin.state() == st
    check_piston(e, twin)
    assert e.read_piston()[0] == 58.0


def chain_bonds(pos, n):
    """Chains of consecutive ids that start within 2 sigma of each other."""
    return [(i, i + 1, (i // 10) % 2) for i in range(n - 1) if (i % 10) != 9 and np.hypot(*((pos[i + 1] - pos[i]) / ONE)) < 2.0]


@pytest.mark.parametrize("mode", MODES)
def test_bonds_match_and_reverse(mode):
    sys, st = walled_fluid(40, 40, 600, t=1.0)
    bonds = chain_bonds(st.pos, st.n())
    assert len(bonds) > 300
    sys = sys.with_bonds([rh.BondType.harmonic(40.0, 1.0), rh.BondType.fene_bond(30.0, 2.0)], bonds)
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(300)
    twin.run(300)
    assert e.read_state() == twin.state()
    e.reverse()
    e.step(300)
    e.reverse()
    assert e.read_state() == st
    kg, peg = e.kicks()
    k, pe, _, _ = O.forces2(twin.sys, O.state2(twin.sys), st.pos, st.species)
    assert np.array_equal(kg, k)
    assert close(peg, pe, 1e-6 * abs(pe))


def test_bonds_added_and_removed_at_run_time():
    sys, st = walled_fluid(40, 40, 600, t=1.0)
    bonds = [(a, b, 0) for a, b, _ in chain_bonds(st.pos, st.n())]
    sys = sys.with_bonds([rh.BondType.harmonic(40.0, 1.0)], bonds)
    e, twin = engine(sys, st, "verlet"), Twin(sys, st)
    e.step(100)
    twin.run(100)
    a, b = bonds[0][0], bonds[0][1]
    assert e.remove_bond(b, a) and not e.remove_bond(b, a)
    twin.update(sys.with_bonds(sys.bond_types, bonds[1:]))
    e.step(100)
    twin.run(100)
    assert e.read_state() == twin.state()
    assert e.add_bond(a, b, 0) and not e.add_bond(a, b, 0)
    twin.update(sys.with_bonds(sys.bond_types, bonds[1:] + [(a, b, 0)]))
    e.step(100)
    twin.run(100)
    assert e.read_state() == twin.state()


def test_kicks_with_every_boundary_term():
    sys, st = walled_fluid(64, 40, 900, t=1.5, fill_y=32)
    bonds = chain_bonds(st.pos, st.n())
    sys = (sys.with_compartments(0, 4).with_walls(0, rh.Wall.Soft, rh.Wall.Soft).with_walls(1, rh.Wall.Soft, rh.Wall.Hard)
           .with_piston(31.0, 10.0, speed=-0.2).with_bonds([rh.BondType.fene_bond(30.0, 2.5)], [(a, b, 0) for a, b, _ in bonds]))
    for mode in MODES:
        e = engine(sys, st, mode)
        kg, peg = e.kicks()
        o = oracle_system(sys)
        k, pe, _, _ = O.forces2(o, O.state2(o), st.pos, st.species)
        assert np.array_equal(kg, k)
        assert close(peg, pe, 1e-6 * abs(pe))
        assert close(e.read_potential(), pe, 1e-6 * abs(pe))


def test_bit_flips():
    sys, st = walled_fluid(40, 40, 600, t=1.0)
    sys = sys.with_compartments(0, 3)  # the boundary path
    e = engine(sys, st, "verlet")
    e.step(50)
    p0 = e.read_positions()
    e.flip_bit(17, 0, 5)
    p1 = e.read_positions()
    want = p0.copy()
    want[17, 0] ^= 1 << 5
    assert np.array_equal(p1, want)
    e.flip_species_bit(1, 1, 3)
    want[:, 1] ^= 1 << 3
    assert np.array_equal(e.read_positions(), want)
    with pytest.raises(rh.RahmanError):
        e.flip_bit(3, 0, 45)  # 2^13 sigma: outside the box
    # any of the 64 bits is taken (engine.rs:964-973); from bit 52 up every flip leaves the box
    for bit in (52, 62, 63):
        with pytest.raises(rh.RahmanError, match="outside the box"):
            e.flip_bit(3, 1, bit)
    with pytest.raises(rh.RahmanError):
        e.flip_bit(3, 0, 64)
    assert np.array_equal(e.read_positions(), want)
    # the run continues exactly like the oracle's from the flipped state
    s = e.read_state()
    twin = Twin(sys, s)
    twin.st.clock = 50
    e.step(100)
    twin.run(100)
    assert e.read_state() == twin.state()


@pytest.mark.parametrize("mode", MODES)
def test_update_system_switches_wall_kinds(mode):
    """Hard walls (the plain path) become heat-bath walls (the boundary path) and back, as
    GpuEngine::update_system allows (engine.rs:895-934); the ledger carries over."""
    sys, st = walled_fluid(40, 40, 600, t=0.5)
    hot = sys.with_walls(1, rh.thermal_wall(2.0), rh.Wall.Hard)
    e, twin = engine(sys, st, mode), Twin(sys, st)
    e.step(100)
    twin.run(100)
    e.update_system(hot)
    twin.update(hot)
    e.step(200)
    twin.run(200)
    assert e.read_state() == twin.state()
    e.update_system(sys)
    twin.update(sys)
    e.step(100)
    twin.run(100)
    assert e.read_state() == twin.state()
    check_ledger(e, twin)


def test_region_and_species_commands():
    sys, st = walled_fluid(40, 40, 600, t=1.0, masses=(1.0, 1.0))
    sys = sys.with_membrane(0, 20.0, [])
    e = engine(sys, st, "verlet")
    e.set_species([1, 2, 3], 1)
    e.set_species_velocity(1, [0.5, -0.25])
    s = e.read_state()
    assert list(s.species[[1, 2, 3]]) == [1, 1, 1]
    assert (s.vel[[1, 2, 3]] == [sys.to_step_displacement(0.5), sys.to_step_displacement(-0.25)]).all()
    e.kick_region([0.0, 0.0], [20.0, 40.0], 0.0, [0.0, 0.0])
    s2 = e.read_state()
    left = s2.pos[:, 0] < 20 * ONE
    assert (s2.vel[left] == 0).all() and np.array_equal(s2.vel[~left], s.vel[~left])
