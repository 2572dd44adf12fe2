"""The batched engine: every system in a mixed batch (2D and 3D, small and large, single-cell
boxes, mixtures) is bit-identical to the oracle and to its own solo run, with and without
hipGraph chunks, and when systems join and leave mid-run."""
from dataclasses import replace

import numpy as np
import pytest

import micrographia_amd as rh
from micrographia_amd.batch import BatchEngine
from micrographia_amd.workloads import ensemble, make_system
from oracle_twin import oracle_state

pytestmark = pytest.mark.gpu

GB = 1 << 30


def mixed_systems():
    out = [make_system(3, 256, 0.9, 1.5, 11),      # box 6.6 sigma: one cell per axis
           make_system(3, 300, 0.3, 2.0, 12),
           make_system(2, 256, 0.9, 0.5, 13),
           make_system(3, 1500, 0.8442, 1.2, 14),
           make_system(2, 5000, 0.7, 1.0, 15),
           make_system(3, 6000, 0.5, 0.8, 16),
           make_system(2, 20000, 0.3, 2.0, 17),
           make_system(3, 9000, 0.9, 0.6, 18)]
    # a binary mixture with pinned walls of species 2 (3D) and a hard-walled 2D box under gravity
    sys, st = make_system(3, 3000, 0.7, 1.0, 19)
    sys = rh.System3.new(sys.box_len, 0.005, rh.Pair.lj(2.5), [1.0, 3.0]).with_pair(0, 1, rh.Pair(0.5, 1.2, 3.0))
    st.species[::3] = 1
    st.vel = rh.maxwell(sys, st.species, 1.0, 19)
    out.append((sys, st))
    sys2 = rh.System.new(rh.box_for(3000, 0.4, 1.0), 0.005, rh.Pair.lj(2.5), [1.0]).with_hard_walls(0).with_hard_walls(1).with_gravity((0.0, -0.05))
    out.append((sys2, rh.fluid(sys2, 3000, 1.0, 20)))
    return out


@pytest.mark.parametrize("graph,tiles", [(0, None), (50, None), (50, (0, 1.5)), (0, (1, 1.25)), (50, (1, 1.0, 512)), (0, (1, 1.25, 1024)),
                                         (0, (2, 1.0)), (50, (2, 1.25, 512))])
def test_each_system_matches_the_oracle(graph, tiles):
    """Every system's trajectory equals the oracle's, whatever the tile settings (list variant:
    the half list built over the row-ordered copy (1, in 3D) or over the cell stencil (2), or
    every pair at both ends (0); cell occupancy and half-list block, as the autotuner installs
    them per size bucket)."""
    systems = mixed_systems()
    b = BatchEngine(budget_bytes=2 * GB, max_systems=64, graph_chunk=graph)
    if tiles is not None:
        for dim in (2, 3):
            for cls in range(8, 16):
                b.set_tuning(dim, cls, *tiles)
    ids = [b.add(sys, st, 150, skin=0.3) for sys, st in systems]
    assert all(i >= 0 for i in ids)
    b.step_all(150)
    left = b.poll()
    assert all(v == 0 for v in left.values())
    for i, (sys, st) in zip(ids, systems):
        got = b.read_state(i)
        want = oracle_state(sys, st, 150)
        assert got == want, f"system {i} ({type(sys).__name__}, n={st.n()}) differs from the oracle"
        assert b.list_stats(i).steps_done == 150


def stochastic_systems():
    """3D systems whose steps draw random numbers: the Andersen bath, heat-bath walls under
    gravity, and a mixture with the bath on one species only."""
    sys, st = make_system(3, 2000, 0.6, 1.0, 21)
    bath = (sys.with_thermostat(1.6, 2.0).with_seed(77), st)
    sys, st = make_system(3, 2000, 0.3, 1.5, 22)
    sys = sys.with_walls(2, rh.thermal_wall(2.0), rh.thermal_wall(0.5)).with_gravity((0.0, 0.0, -0.2)).with_seed(78)
    st.pos[:, 2] = np.clip(st.pos[:, 2], (1 << 32) // 2, sys.box_len[2] - (1 << 32) // 2)
    walls = (sys, st)
    sys, st = make_system(3, 3000, 0.7, 1.0, 23)
    sys = rh.System3.new(sys.box_len, 0.005, rh.Pair.lj(2.5), [1.0, 3.0]).with_thermostat(0.8, 4.0).with_seed(79)
    sys = replace(sys, thermostat=replace(sys.thermostat, species=0b10))
    st.species[::3] = 1
    st.vel = rh.maxwell(sys, st.species, 1.0, 23)
    mix = (sys, st)
    return [bath, walls, mix]


@pytest.mark.parametrize("graph,block", [(0, 256), (50, 256), (50, 1024)])
def test_bath_and_heat_bath_walls_match_the_oracle(graph, block):
    """The Andersen bath and heat-bath walls draw the same Threefry numbers as md-core
    (three.rs:232-265), keyed by particle id and the system's own clock, so each system's
    trajectory equals the oracle's bit for bit beside systems without them."""
    stoch = stochastic_systems()
    plain = [make_system(3, 1500, 0.8442, 1.2, 24), make_system(2, 3000, 0.7, 1.0, 25)]
    b = BatchEngine(budget_bytes=2 * GB, max_systems=16, graph_chunk=graph)
    for dim in (2, 3):
        for cls in range(8, 16):
            b.set_tuning(dim, cls, 1, 1.0, block)
    steps = [150, 300, 200, 150, 150]
    systems = stoch + plain
    ids = [b.add(sys, st, k, skin=0.3) for (sys, st), k in zip(systems, steps)]
    assert all(i >= 0 for i in ids)
    b.step_all(max(steps))
    for i, (sys, st), k in zip(ids, systems, steps):
        want = oracle_state(sys, st, k)
        assert b.read_state(i) == want, f"system {i} (n={st.n()}) differs from the oracle"
    # the draws took effect: without the bath, or with hard walls in place of heat-bath ones,
    # the trajectories differ
    (s0, st0), (s1, st1) = stoch[0], stoch[1]
    assert b.read_state(ids[0]) != oracle_state(replace(s0, thermostat=None), st0, steps[0])
    hard = s1.with_walls(2, rh.Wall.Hard, rh.Wall.Hard)
    assert b.read_state(ids[1]) != oracle_state(hard, st1, steps[1])


ONE = 1 << 32


def squeeze(st, box_len, axis):
    """Move the start off the walls of `axis`: positions scaled into [1, L - 1] sigma."""
    L = int(box_len[axis])
    st.pos[:, axis] = ONE + (st.pos[:, axis].astype(np.float64) * (L - 2 * ONE) / L).astype(np.int64)


def booked_systems():
    """Systems whose steps book the ledger: a 3D liquid between hard and soft walls on two axes,
    a 2D gas between soft walls and hard walls under gravity, the stochastic 3D systems, a 2D
    gas between heat-bath walls at 2.0 and 0.5 under gravity, and 2D thermostat regions: a fan
    (a slab along x with a drift) and a bath layer (a slab along y)."""
    sys, st = make_system(3, 2000, 0.5, 1.5, 31)
    sys = sys.with_walls(2, rh.Wall.Soft, rh.Wall.Hard).with_walls(0, rh.Wall.Hard, rh.Wall.Soft)
    squeeze(st, sys.box_len, 0)
    squeeze(st, sys.box_len, 2)
    walls3 = (sys, st)
    sys, st = make_system(2, 3000, 0.5, 1.2, 32)
    sys = sys.with_walls(0, rh.Wall.Soft, rh.Wall.Soft).with_hard_walls(1).with_gravity((0.0, -0.05))
    squeeze(st, sys.box_len, 0)
    squeeze(st, sys.box_len, 1)
    walls2 = (sys, st)
    sys, st = make_system(2, 2500, 0.3, 1.0, 35)
    sys = sys.with_walls(1, rh.thermal_wall(2.0), rh.thermal_wall(0.5)).with_gravity((0.0, -0.1)).with_seed(80)
    squeeze(st, sys.box_len, 1)
    heat2 = (sys, st)
    sys, st = make_system(2, 3000, 0.6, 0.6, 36)
    fan = (sys.with_thermostat(1.5, 4.0).with_fan(0b1, 2.0, 20.0, 0.3).with_seed(81), st)  # species mask
    sys, st = make_system(2, 3000, 0.6, 0.6, 37)
    layer = (sys.with_thermostat(1.8, 3.0).with_bath_layer(0.0, 30.0).with_seed(82), st)
    return [walls3, walls2] + stochastic_systems() + [heat2, fan, layer]


@pytest.mark.parametrize("graph", [0, 50])
def test_ledgers_equal_the_systems_run_alone(graph):
    """Each system books its wall and bath exchanges in its own ledger with the single-system
    engine's fixed-point, order-independent sums, so the batched ledger equals the one the
    system books alone, term for term, beside systems that book nothing; the trajectories equal
    the oracle's."""
    booked = booked_systems()
    plain = [make_system(3, 1500, 0.8442, 1.2, 33), make_system(2, 3000, 0.7, 1.0, 34)]
    b = BatchEngine(budget_bytes=2 * GB, max_systems=16, graph_chunk=graph)
    steps = [300, 300, 150, 300, 200, 300, 200, 200, 150, 150]
    systems = booked + plain
    ids = [b.add(sys, st, k, skin=0.3) for (sys, st), k in zip(systems, steps)]
    assert all(i >= 0 for i in ids)
    b.step_all(max(steps))
    for i, (sys, st), k in zip(ids, systems, steps):
        assert b.read_state(i) == oracle_state(sys, st, k), f"system {i} (n={st.n()}) differs from the oracle"
        if isinstance(sys, rh.System3):
            solo = rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, st.n(), 0.3))
        else:
            solo = rh.GpuEngine(sys, st, rh.verlet_with_skin(sys, st.n(), 0.3))
        solo.step(k)
        got, want = b.read_ledger(i), solo.read_ledger()
        solo.close()
        for field in ("energy", "impulse", "hits"):
            assert np.array_equal(getattr(got, field), getattr(want, field)), f"system {i}: ledger {field} {getattr(got, field)} vs {getattr(want, field)}"
    # the walls and the bath were hit and booked; the periodic systems booked nothing
    led3 = b.read_ledger(ids[0])  # hard walls are crossed (hits), soft ones push back (impulse)
    assert led3.hits[[0, 5]].min() > 0 and np.count_nonzero(led3.impulse[[1, 4]]) == 2
    led2 = b.read_ledger(ids[1])
    assert led2.hits[:4].sum() > 0 and np.count_nonzero(led2.impulse[:4]) >= 2
    assert b.read_ledger(ids[2]).hits[6] > 0 and b.read_ledger(ids[3]).hits[4:6].min() > 0
    assert b.read_ledger(ids[5]).hits[2:4].min() > 0  # the 2D heat-bath walls
    assert b.read_ledger(ids[6]).hits[6] > 0 and b.read_ledger(ids[7]).hits[6] > 0  # the 2D thermostat regions
    for i in ids[-2:]:
        assert not b.read_ledger(i).hits.any()


def test_alone_and_in_the_mix_agree():
    specs = ensemble(24, meta_seed=99, n_hi=20000)
    built = [s.build() for s in specs]
    b = BatchEngine(budget_bytes=3 * GB, max_systems=64, graph_chunk=100)
    ids = [b.add(sys, st, 300) for sys, st in built]
    b.step_all(300)
    for i, (sys, st) in zip(ids, built):
        if isinstance(sys, rh.System3):
            e = rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, st.n(), 0.3))
        else:
            e = rh.GpuEngine(sys, st, rh.verlet_with_skin(sys, st.n(), 0.3))
        e.step(300)
        assert b.read_state(i) == e.read_state()


def test_join_and_leave_mid_run():
    systems = mixed_systems()[:6]
    b = BatchEngine(budget_bytes=2 * GB, max_systems=16, graph_chunk=40)
    first = [b.add(sys, st, 200) for sys, st in systems[:3]]
    b.run(80)
    b.poll()
    b.remove(first[1])                       # leaves after 80 steps
    late = [b.add(sys, st, 120) for sys, st in systems[3:]]
    b.run(130)
    left = b.poll()
    assert all(v == 0 for v in left.values())  # 210 iterations cover 200 and 80 + 120
    assert b.read_state(first[0]) == oracle_state(*systems[0], 200)
    assert b.read_state(first[2]) == oracle_state(*systems[2], 200)
    for i, (sys, st) in zip(late, systems[3:]):
        assert b.read_state(i) == oracle_state(sys, st, 120)


def test_the_full_512_system_ensemble_matches_the_oracle():
    """The benchmark's own ensemble at full shape: 512 systems, 2D and 3D 50/50, N log-uniform
    in [256, 65536] (5.9M atoms), density U[0.3, 0.9], T U[0.5, 2.0], meta-seed 2026, advanced
    through two replays of a captured three-step graph; every system equals the oracle's
    IntEngine<f32> / IntEngine3<f32> bit for bit."""
    import concurrent.futures as cf

    specs = ensemble(512, meta_seed=2026)
    built = [s.build() for s in specs]
    assert sum(st.n() for _, st in built) > 5_000_000 and max(st.n() for _, st in built) > 60_000
    b = BatchEngine(budget_bytes=12 * GB, max_systems=1024, graph_chunk=3)
    ids = [b.add(sys, st, 6) for sys, st in built]
    assert all(i >= 0 for i in ids)
    b.step_all(6)
    assert all(v == 0 for v in b.poll().values())
    # the oracle releases the interpreter inside its C calls, so threads run it side by side
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        want = list(ex.map(lambda sb: oracle_state(sb[0], sb[1], 6), built))
    bad = [k for k, (i, w) in enumerate(zip(ids, want)) if b.read_state(i) != w]
    assert not bad, f"systems {bad[:10]} differ from the oracle"


@pytest.mark.parametrize("variant,density", [(0, 1.6), (1, 2.2)])
def test_a_list_that_overflows_regrows_and_stays_exact(variant, density):
    """A dense cluster (FCC) in a dilute box (mean density 0.15): the capacity rule sizes the
    list for the box's density, so the first build overflows (every pair at both ends at cluster
    density 1.6: ~147 neighbours against a capacity of 120; the half list built over the rows at
    2.2: ~160 entries).  The system stops, poll() regrows its list, step_all() runs the steps
    still left, and the trajectory equals the oracle's, beside a system that never overflowed."""
    n = 2000
    small = rh.cube_for(n, density)
    big = rh.cube_for(n, 0.15)
    pos = rh.fcc(n, small, rh.fcc_cells(n, small), 0.02, 5)
    sys = rh.System3.new(big, 0.002, rh.Pair.lj(2.5), [1.0])
    sp = np.zeros(n, np.uint32)
    st = rh.State3(pos, rh.maxwell(sys, sp, 0.5, 5), sp)
    other = make_system(3, 3000, 0.8, 1.0, 6)
    b = BatchEngine(budget_bytes=2 * GB, max_systems=8, graph_chunk=0)
    b.set_tuning(3, 10, variant, 1.0)
    i = b.add(sys, st, 60, skin=0.3)
    cap0 = b.list_stats(i).cap
    j = b.add(*other, steps=60)
    b.step_all(60)
    assert b.list_stats(i).overflowed and not b.list_stats(j).overflowed
    assert b.list_stats(i).cap > cap0 and b.list_stats(i).steps_done == 60
    assert b.read_state(i) == oracle_state(sys, st, 60)
    assert b.read_state(j) == oracle_state(*other, 60)


def test_admission_refuses_what_does_not_fit():
    b = BatchEngine(budget_bytes=64 << 20, max_systems=8, graph_chunk=0)
    small = make_system(3, 2000, 0.8, 1.0, 3)
    big = make_system(3, 200000, 0.8, 1.0, 4)
    assert b.bytes_needed(big[0], 200000) > 64 << 20
    assert b.add(*big, steps=10) == -1
    i = b.add(*small, steps=10)
    assert i >= 0
    b.step_all(10)
    assert b.read_state(i) == oracle_state(*small, 10)
