"""The 3D engine against the oracle (IntEngine3<f32> bit for bit, IntEngine3<f64> within the
reference's tolerances) and its exactness properties, mirroring crates/md/tests/three.rs."""
import math

import numpy as np
import pytest

import micrographia_amd as rh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ONE = 1 << 32


def liquid(n, density, t, walls=False, seed_pos=7, seed_vel=3):
    """crates/md/tests/three.rs:15-31 (product md-core; equal to the oracle's, see test_host)."""
    sys = rh.System3.new(rh.cube_for(n, density), 0.005, rh.Pair.lj(2.5), [1.0])
    if walls:
        sys = sys.with_walls(2, rh.Wall.Hard, rh.Wall.Hard)
    pos = rh.fcc(n, sys.box_len, rh.fcc_cells(n, sys.box_len), 0.05, seed_pos)
    if walls:
        pos[:, 2] = np.clip(pos[:, 2], ONE // 2, sys.box_len[2] - ONE // 2)
    sp = np.zeros(n, np.uint32)
    vel = rh.maxwell(sys, sp, t, seed_vel)
    return sys, rh.State3(pos, vel, sp)


def oracle_system(sys: rh.System3) -> O.System3:
    o = O.System3()
    c = sys.to_c()
    for a in range(3):
        o.box_len[a] = c.box_len[a]
        o.walled[a] = c.walled[a]
        o.gravity[a] = c.gravity[a]
        for s in range(2):
            o.wall_kind[a][s] = c.wall_kind[a][s]
            o.wall_t[a][s] = c.wall_t[a][s]
    o.dt = c.dt
    o.nspecies = c.nspecies
    o.pinned = c.pinned
    for k in range(4):
        o.masses[k] = c.masses[k]
    for k in range(16):
        o.pairs[k] = O.Pair(c.pairs[k].eps, c.pairs[k].sigma, c.pairs[k].rc)
    o.seed = c.seed
    o.has_bath = c.has_bath
    o.bath_species = c.bath_species
    o.bath_t = c.bath_t
    o.bath_rate = c.bath_rate
    return o


def oracle_run(sys, st: rh.State3, steps, prec=32):
    pos, vel, sp = st.pos.copy(), st.vel.copy(), st.species.copy()
    O.step3(oracle_system(sys), pos, vel, sp, steps, prec)
    return rh.State3(pos, vel, sp)


MODES = {
    "cells": lambda sys, n: rh.ForceMode.cell_list(),
    "verlet": lambda sys, n: rh.verlet3_with_skin(sys, n, 0.3),
}

# (kernel variant, block): the full list, the half list at each block size (built from the
# row-ordered copy), and the half list built over the cell stencil
VARIANTS = [(0, 256), (1, 256), (1, 512), (1, 1024), (1, 2048), (1, 4096), (2, 512), (2, 2048)]


def engine(sys, st, mode, variant=1, block=512):
    e = rh.GpuEngine3(sys, st, mode)
    if mode.kind == 1:
        e.tune(variant, block)
    return e


@pytest.mark.parametrize("mode", ["cells", "verlet"])
def test_kicks_are_bit_exact_with_the_f32_twin(mode):
    """three.rs:33-64, strengthened: every kick identical to IntEngine3<f32>, rms vs f64 < 1e-5."""
    sys, s0 = liquid(4000, 0.8, 1.0)
    st = oracle_run(sys, s0, 100)
    gpu = rh.GpuEngine3(sys, st, MODES[mode](sys, st.n()))
    kg, peg = gpu.kicks()
    osys = oracle_system(sys)
    kc, pec = O.kicks3(osys, st.pos, st.species, 32)
    k64, pe64 = O.kicks3(osys, st.pos, st.species, 64)
    print(f"identical kicks: {(kg == kc).all(axis=1).sum()}/{len(kg)}")
    assert np.array_equal(kg, kc)
    rel = math.sqrt(((kg - k64).astype(np.float64) ** 2).sum() / (k64.astype(np.float64) ** 2).sum())
    assert rel < 1e-5
    assert abs((peg - pec) / pec) < 1e-5
    assert abs((peg - pe64) / pe64) < 1e-5


@pytest.mark.parametrize("mode", ["cells", "verlet"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_trajectory_is_bit_exact_with_the_f32_twin(mode, variant):
    """three.rs:66-80 asks for 1e-5 sigma after 50 steps; integer state here is identical."""
    if mode == "cells" and variant != VARIANTS[0]:
        pytest.skip("the variant applies to Verlet lists only")
    sys, s0 = liquid(2000, 0.8, 1.0, walls=True)
    gpu = engine(sys, s0, MODES[mode](sys, s0.n()), *variant)
    gpu.step(200)
    want = oracle_run(sys, s0, 200)
    got = gpu.read_state()
    assert got == want
    assert got.hash() == want.hash()


def test_long_cutoff_keeps_the_range_test(block=512):
    """A cutoff of 7.5 sigma: list separations may reach the 8 sigma range of the f32
    separation, so the engine keeps the reference's range test (the full-list kernel, whatever
    variant is asked for); the trajectory matches the twin."""
    sys, s0 = liquid(2000, 0.1, 1.0, walls=True)
    sys = rh.System3.new(sys.box_len, 0.005, rh.Pair.lj(7.5), [1.0]).with_walls(2, rh.Wall.Hard, rh.Wall.Hard)
    gpu = engine(sys, s0, rh.verlet3_with_skin(sys, s0.n(), 0.3), 1, block)
    gpu.step(60)
    assert gpu.read_state() == oracle_run(sys, s0, 60)


def test_heavy_species_match_the_twin():
    """A species of mass 1e8 (kick factor about 1e-3: most of its kicks round to zero) beside
    unit masses, bit for bit with the twin."""
    sys, s0 = liquid(2000, 0.8, 1.0)
    sys = rh.System3.new(sys.box_len, 0.005, rh.Pair.lj(2.5), [1.0, 1e8])
    s0.species[::5] = 1
    s0.vel = rh.maxwell(sys, s0.species, 1.0, 4)
    gpu = engine(sys, s0, rh.verlet3_with_skin(sys, s0.n(), 0.3), 1, 512)
    gpu.step(100)
    assert gpu.read_state() == oracle_run(sys, s0, 100)


@pytest.mark.parametrize("variant", VARIANTS)
def test_periodic_liquid_bit_exact_over_many_rebuilds(variant):
    sys, s0 = liquid(4000, 0.8442, 1.2)
    gpu = engine(sys, s0, rh.verlet3_with_skin(sys, 4000, 0.3), *variant)
    gpu.step(300)
    stats = gpu.read_list_stats()
    assert stats.rebuilds >= 10 and not stats.overflowed
    assert gpu.read_state() == oracle_run(sys, s0, 300)


def test_reversal_restores_the_start_bit_for_bit():
    """three.rs:82-93"""
    sys, s = liquid(4000, 0.8, 1.0, walls=True)
    gpu = rh.GpuEngine3(sys, s, rh.verlet3(sys, s.n()))
    gpu.step(1000)
    gpu.reverse()
    gpu.step(1000)
    gpu.reverse()
    back = gpu.read_state()
    assert back.hash() == s.hash()


def test_energy_is_conserved():
    """three.rs:95-111, energies from the oracle's f64 evaluation of GPU states."""
    sys, s = liquid(8000, 0.8, 1.0)
    n = s.n()
    gpu = rh.GpuEngine3(sys, s, rh.verlet3_with_skin(sys, n, 0.3))
    osys = oracle_system(sys)
    energies = []
    for _ in range(10):
        gpu.step(500)
        st = gpu.read_state()
        ke, pe = O.energy3(osys, st.pos, st.vel, st.species, 64)
        energies.append((ke + pe) / n)
    drift = max(energies[2:]) - min(energies[2:])
    print("energy per particle:", [f"{e:.5f}" for e in energies])
    assert drift < 2e-3
    # relative total-energy tolerance over the run
    assert abs(energies[-1] - energies[2]) / abs(energies[2]) < 1e-3


def test_verlet_lists_and_cell_list_give_the_same_trajectory():
    """three.rs:135-151"""
    sys, s = liquid(6000, 0.8, 1.0, walls=True)
    cells = rh.GpuEngine3(sys, s, rh.ForceMode.cell_list())
    lists = rh.GpuEngine3(sys, s, rh.verlet3(sys, s.n()))
    cells.step(2000)
    lists.step(2000)
    a, b = cells.read_state(), lists.read_state()
    stats = lists.read_list_stats()
    print(f"3D: {stats.rebuilds} list rebuilds in 2000 steps")
    assert not stats.overflowed
    assert stats.rebuilds > 10
    assert a.hash() == b.hash()


def test_run_to_run_identical_and_momentum_exact():
    sys, s = liquid(5000, 0.7, 1.0)
    a = rh.GpuEngine3(sys, s, rh.verlet3_with_skin(sys, 5000, 0.3))
    b = rh.GpuEngine3(sys, s, rh.verlet3_with_skin(sys, 5000, 0.3))
    a.step(700)
    b.step(300)
    b.step(400)
    sa, sb = a.read_state(), b.read_state()
    assert sa != s and sa == sb
    assert (sa.vel.astype(np.int64).sum(0) == 0).all()


def test_storage_permutation_does_not_change_the_trajectory():
    """gpu_engine.rs:107-123 in 3D"""
    sys, s = liquid(3000, 0.8, 1.0)
    perm = np.random.default_rng(4).permutation(3000)
    sp = rh.State3(s.pos[perm], s.vel[perm], s.species[perm])
    a = rh.GpuEngine3(sys, s)
    b = rh.GpuEngine3(sys, sp)
    a.step(400)
    b.step(400)
    ra, rb = a.read_state(), b.read_state()
    assert np.array_equal(ra.pos[perm], rb.pos) and np.array_equal(ra.vel[perm], rb.vel)


@pytest.mark.parametrize("variant", [(0, 256), (1, 512), (1, 2048), (1, 4096)])
def test_a_tiny_list_capacity_regrows_and_stays_exact(variant):
    sys, s = liquid(3000, 0.8442, 1.2)
    gpu = engine(sys, s, rh.ForceMode(1, 0.3, 12), *variant)
    gpu.step(150)
    stats = gpu.read_list_stats()
    assert stats.cap > 12
    assert gpu.read_state() == oracle_run(sys, s, 150)


def test_block_size_switches_mid_run_stay_exact():
    """The half list is rebuilt for each block size (its chunk tags name owners within a block);
    switching between sizes and to the full list mid-run leaves the trajectory the twin's."""
    sys, s0 = liquid(4500, 0.8442, 1.2)
    gpu = engine(sys, s0, rh.verlet3_with_skin(sys, s0.n(), 0.3), 1, 512)
    for variant, block in ((1, 512), (1, 2048), (0, 256), (1, 1024), (1, 4096), (1, 256)):
        gpu.tune(variant, block)
        gpu.step(30)
    assert gpu.read_state() == oracle_run(sys, s0, 180)


def test_a_few_particles_fill_part_of_one_wave():
    """100 particles: one full 64-particle wave of the list build and one partly filled."""
    sys, s0 = liquid(100, 0.1, 1.0)
    gpu = engine(sys, s0, rh.verlet3_with_skin(sys, s0.n(), 0.3), 1, 256)
    gpu.step(300)
    assert gpu.read_state() == oracle_run(sys, s0, 300)


def test_mixture_pinned_gravity_soft_walls_match_the_twin():
    n = 3000
    sys, s = liquid(n, 0.7, 1.0, walls=True)
    sys = rh.System3.new(sys.box_len, 0.005, rh.Pair.lj(2.5), [1.0, 3.0, 1.0])
    sys = sys.with_pair(0, 1, rh.Pair(0.5, 1.2, 3.0)).with_pair(1, 2, rh.Pair(1.0, 0.9, 2.2))
    sys = sys.with_walls(2, rh.Wall.Soft, rh.Wall.Hard).with_gravity((0.0, 0.0, -0.1)).with_pinned(2)
    s.species[::3] = 1
    s.species[::7] = 2
    s.pos[:, 2] = np.clip(s.pos[:, 2], ONE, sys.box_len[2] - ONE)
    s.vel = rh.maxwell(sys, s.species, 1.0, 5)
    want = oracle_run(sys, s, 150)
    for mode, variant in ((rh.ForceMode.cell_list(), (0, 256)), (rh.verlet3_with_skin(sys, n, 0.3), (0, 256)),
                          (rh.verlet3_with_skin(sys, n, 0.3), (1, 256)), (rh.verlet3_with_skin(sys, n, 0.3), (1, 1024)),
                          (rh.verlet3_with_skin(sys, n, 0.3), (1, 2048))):
        gpu = engine(sys, s, mode, *variant)
        gpu.step(150)
        assert gpu.read_state() == want


def test_set_species_changes_the_chosen_particles():
    """three.rs:153-166"""
    sys, s = liquid(2000, 0.8, 1.0)
    sys = rh.System3.new(sys.box_len, 0.005, rh.Pair.lj(2.5), [1.0, 1.0])
    gpu = rh.GpuEngine3(sys, s)
    gpu.step(300)
    gpu.set_species([5, 17, 1999], 1)
    gpu.step(2000)
    st = gpu.read_state()
    assert list(np.nonzero(st.species == 1)[0]) == [5, 17, 1999]


def test_upload_round_trips_and_kick_region():
    sys, s = liquid(2000, 0.8, 1.0)
    gpu = rh.GpuEngine3(sys, s)
    gpu.step(50)
    mid = gpu.read_state()
    gpu.upload(s)
    assert gpu.read_state() == s
    gpu.upload(mid)
    assert gpu.read_state() == mid
    b = sys.box_f64()
    gpu.kick_region([0, 0, 0], [b[0] / 2, b[1], b[2]], 0.0, [0.0, 0.0, 0.0])
    st = gpu.read_state()
    x = mid.pos[:, 0] / ONE
    left, right = x < b[0] / 2 - 1e-4, x >= b[0] / 2 + 1e-4
    assert (st.vel[left] == 0).all() and (st.vel[right] == mid.vel[right]).all()
    assert left.sum() > 500 and right.sum() > 500


def test_kinetic_energy_readout():
    sys, s = liquid(4000, 0.8, 1.0)
    gpu = rh.GpuEngine3(sys, s)
    assert abs(gpu.read_kinetic_total() - s.half_step_kinetic(sys)) < 1e-9 * s.half_step_kinetic(sys)


def test_planned_memory_equals_what_the_engines_allocate():
    """bytes_needed() (a host-side plan, the serving loop's admission figure) equals the device
    memory each engine then holds: 3D in both force modes and both list builds' geometries, 2D
    plain and with boundaries."""
    from micrographia_amd.engine import bytes_needed

    for n in (4000, 262144):
        sys, s = liquid(n, 0.8442, 1.0)
        for mode in (rh.verlet3_with_skin(sys, n, 0.3), rh.ForceMode.cell_list()):
            e = rh.GpuEngine3(sys, s, mode)
            assert e.device_bytes() == bytes_needed(sys, n, mode), (n, mode)
            e.close()
    sys2 = rh.System.new((40 << 32, 40 << 32), 0.005, rh.Pair.lj(2.5), [1.0])
    walled = sys2.with_hard_walls(0).with_hard_walls(1).with_compartments(0, 3)  # five 8-sigma boxes
    for sy in (sys2, walled):
        st2 = rh.fluid(sy, 600, 1.0, 3)
        mode = rh.verlet_with_skin(sy, 600, 0.3)
        e = rh.GpuEngine(sy, st2, mode)
        assert e.device_bytes() == bytes_needed(sy, 600, mode)
        e.close()


def test_per_particle_kinetic_energies_in_id_order():
    """read_kinetic (engine3.rs:626-631): 1/2 m v^2 per particle from the half-step velocity,
    zero for pinned species, with the species, in particle-id order however storage is sorted."""
    _, s = liquid(3000, 0.8, 1.2)
    n = s.n()
    sp = (np.arange(n) % 3).astype(np.uint32)
    base, _ = liquid(3000, 0.8, 1.2)
    sys = rh.System3.new(base.box_len, 0.005, rh.Pair.lj(2.5), [1.0, 2.0, 0.5]).with_pinned(2)
    s = rh.State3(s.pos, s.vel, sp)
    gpu = rh.GpuEngine3(sys, s, rh.verlet3_with_skin(sys, n, 0.3))
    gpu.step(40)  # storage is Morton-sorted by now
    st = gpu.read_state()
    ke, species = gpu.read_kinetic()
    assert ke.dtype == np.float32 and species.dtype == np.uint32
    assert np.array_equal(species, sp)
    m = np.array([1.0, 2.0, 0.5])[sp]
    v = st.vel.astype(np.float64) / 2.0**32 / sys.dt
    want = 0.5 * m * np.sum(v * v, axis=1)
    want[sp == 2] = 0.0
    # one f32 rounding of the f64 value
    assert np.array_equal(ke, want.astype(np.float32))
    assert np.all(ke[sp == 2] == 0.0)
    gpu.measure_kinetic()
    assert abs(gpu.read_kinetic_total() - want.sum()) < 1e-9 * want.sum()


def test_benchmark_size_short_run_bit_exact():
    """262,144 atoms (the single-GPU benchmark's liquid) for a few steps against the oracle."""
    n = 262144
    sys = rh.System3.new(rh.cube_for(n, 0.8442), 0.005, rh.Pair.lj(2.5), [1.0])
    pos = rh.fcc(n, sys.box_len, rh.fcc_cells(n, sys.box_len), 0.05, 1)
    sp = np.zeros(n, np.uint32)
    s = rh.State3(pos, rh.maxwell(sys, sp, 0.72, 1), sp)
    gpu = rh.GpuEngine3(sys, s, rh.verlet3_with_skin(sys, n, 0.3))
    gpu.step(3)
    assert gpu.read_state() == oracle_run(sys, s, 3)


def test_large_system_properties():
    """4,194,304 atoms at T = 1.2: zero integer momentum and exact reversal at full size."""
    n = 4194304
    sys = rh.System3.new(rh.cube_for(n, 0.8442), 0.005, rh.Pair.lj(2.5), [1.0])
    pos = rh.fcc(n, sys.box_len, rh.fcc_cells(n, sys.box_len), 0.05, 2)
    sp = np.zeros(n, np.uint32)
    s = rh.State3(pos, rh.maxwell(sys, sp, 1.2, 2), sp)
    gpu = rh.GpuEngine3(sys, s, rh.verlet3_with_skin(sys, n, 0.3))
    gpu.step(60)
    mid = gpu.read_state()
    assert (mid.vel.astype(np.int64).sum(0) == 0).all()
    stats = gpu.read_list_stats()
    assert stats.rebuilds >= 3 and not stats.overflowed
    gpu.reverse()
    gpu.step(60)
    gpu.reverse()
    assert gpu.read_state().hash() == s.hash()


# ---- stochastic boundaries and the ledger (int_engine3.rs:138-206, three.rs:232-265) --------
# The draws are f64 on both sides (the device's log/cos/sin against the C library's); a
# velocity is rounded to an integer displacement, so the two agree unless a one-ulp difference
# straddles a rounding boundary (about 1e-8 per draw).  The ledger sums fixed-point terms of
# 2^-32 on the device, so its totals differ from the twin's f64 sums by at most that rounding
# per term.

def oracle_run_ledger(sys, st, steps, prec=32):
    pos, vel, sp = st.pos.copy(), st.vel.copy(), st.species.copy()
    _, led = O.step3(oracle_system(sys), pos, vel, sp, steps, prec)
    return rh.State3(pos, vel, sp), led


def ledger_close(got, want, walls):
    for w in walls:
        assert int(got.hits[w]) == int(want.hits[w]), f"hits of wall {w}"
        n = max(1, int(want.hits[w]))
        for field in ("energy", "impulse"):
            g, x = getattr(got, field)[w], getattr(want, field)[w]
            assert abs(g - x) <= n * 2.0 ** -32 + 1e-9 * abs(x) + 1e-12, f"{field} of wall {w}: {g} vs {x}"


@pytest.mark.parametrize("mode", ["cells", "verlet"])
def test_andersen_bath_matches_the_twin(mode):
    sys, st = liquid(2000, 0.6, 1.0)
    sys = sys.with_thermostat(1.6, 2.0)
    want, led = oracle_run_ledger(sys, st, 150)
    assert led.hits[6] > 1000
    gpu = engine(sys, st, MODES[mode](sys, st.n()))
    gpu.step(150)
    assert gpu.read_state() == want
    ledger_close(gpu.read_ledger(), led, [6])


@pytest.mark.parametrize("mode", ["cells", "verlet"])
def test_heat_bath_walls_match_the_twin(mode):
    sys, st = liquid(2000, 0.3, 1.5, walls=True)
    sys = sys.with_walls(2, rh.thermal_wall(2.0), rh.thermal_wall(0.5)).with_gravity((0.0, 0.0, -0.2))
    want, led = oracle_run_ledger(sys, st, 300)
    assert led.hits[4] + led.hits[5] > 20
    gpu = engine(sys, st, MODES[mode](sys, st.n()))
    gpu.step(300)
    assert gpu.read_state() == want
    ledger_close(gpu.read_ledger(), led, [4, 5])


def squeeze(st, sys, axis):
    """Move the lattice off the walls of `axis`: positions scaled into [1, L - 1] sigma."""
    L = int(sys.box_len[axis])
    st.pos[:, axis] = ONE + (st.pos[:, axis].astype(np.float64) * (L - 2 * ONE) / L).astype(np.int64)


def test_hard_and_soft_wall_ledger_matches_the_twin():
    sys, st = liquid(2000, 0.5, 1.5)
    sys = sys.with_walls(2, rh.Wall.Soft, rh.Wall.Hard).with_walls(0, rh.Wall.Hard, rh.Wall.Soft)
    squeeze(st, sys, 0)
    squeeze(st, sys, 2)
    want, led = oracle_run_ledger(sys, st, 300)
    assert min(led.hits[0], led.hits[5]) > 0 and np.isfinite(led.impulse).all()
    for variant in VARIANTS[:2] + VARIANTS[-1:]:
        gpu = engine(sys, st, rh.verlet3_with_skin(sys, st.n(), 0.3), *variant)
        gpu.step(300)
        assert gpu.read_state() == want
        ledger_close(gpu.read_ledger(), led, [0, 1, 4, 5])


def test_bath_sets_the_temperature():
    """crates/md/tests/three.rs:114-121: an Andersen bath at 1.6 brings a cold liquid to kT 1.6."""
    sys, st = liquid(4000, 0.6, 0.5)
    sys = sys.with_thermostat(1.6, 2.0)
    gpu = engine(sys, st, rh.verlet3_with_skin(sys, st.n(), 0.3))
    gpu.step(1500)
    t = 2.0 * gpu.read_kinetic_total() / (3 * st.n())
    assert abs(t - 1.6) < 0.1, t
