"""The 2D engine against the oracle's IntEngine<f32> (crates/md/tests/gpu_engine.rs)."""
import math

import numpy as np
import pytest

import micrographia_amd as rh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fluid(n, density, pair, seed):
    sys = rh.System.new(rh.box_for(n, density, 1.0), 0.005, pair, [1.0])
    return sys, rh.fluid(sys, n, 1.0, seed)


def oracle_system(sys: rh.System) -> O.System2:
    o = O.System2()
    c = sys.to_c()
    for a in range(2):
        o.box_len[a] = c.box_len[a]
        o.walled[a] = c.walled[a]
        o.gravity[a] = c.gravity[a]
    o.dt = c.dt
    o.nspecies = c.nspecies
    o.pinned = c.pinned
    for k in range(4):
        o.masses[k] = c.masses[k]
    for k in range(16):
        o.pairs[k] = O.Pair(c.pairs[k].eps, c.pairs[k].sigma, c.pairs[k].rc)
    return o


def oracle_run(sys, st, steps, prec=32):
    pos, vel, sp = st.pos.copy(), st.vel.copy(), st.species.copy()
    O.step2(oracle_system(sys), pos, vel, sp, steps, prec)
    return rh.State(pos, vel, sp)


@pytest.mark.parametrize("pair", [rh.Pair.wca(), rh.Pair.lj(2.5)])
def test_kicks_match_the_twin(pair):
    """gpu_engine.rs:27-60, strengthened to identical kicks."""
    sys, s0 = fluid(3000, 0.7, pair, 1)
    st = oracle_run(sys, s0, 200)
    gpu = rh.GpuEngine(sys, st)
    kg, peg = gpu.kicks()
    osys = oracle_system(sys)
    k32, pe32 = O.kicks2(osys, st.pos, st.species, 32)
    k64, _ = O.kicks2(osys, st.pos, st.species, 64)
    assert np.array_equal(kg, k32)
    rel = math.sqrt(((kg - k64).astype(np.float64) ** 2).sum() / (k64.astype(np.float64) ** 2).sum())
    assert rel < 1e-5
    assert abs((peg - pe32) / pe32) < 1e-5


@pytest.mark.parametrize("cells,block", [(True, None), (False, None), (False, 256), (False, 1024), (False, 2048)])
def test_trajectory_matches_the_twin(cells, block):
    """Config 1's engine (2D LJ, rho 0.7, 64x64 lattice) for 300 steps, bit for bit, with the
    half list in work-groups of each size."""
    sys, s0 = fluid(4096, 0.7, rh.Pair.lj(2.5), 2)
    mode = rh.ForceMode.cell_list() if cells else rh.verlet_for(sys, s0.n())
    gpu = rh.GpuEngine(sys, s0, mode)
    if block is not None:
        gpu.tune(1, block)
    gpu.step(300)
    assert gpu.read_state() == oracle_run(sys, s0, 300)


def test_reversal_is_exact_and_run_to_run_identical():
    sys, s0 = fluid(5000, 0.6, rh.Pair.wca(), 3)
    for mode in (rh.ForceMode.cell_list(), rh.verlet_for(sys, s0.n())):
        e = rh.GpuEngine(sys, s0, mode)
        e.step(1000)
        e.reverse()
        e.step(1000)
        e.reverse()
        assert e.read_state() == s0
    a = rh.GpuEngine(sys, s0)
    b = rh.GpuEngine(sys, s0)
    a.step(500)
    b.step(500)
    assert a.read_state() == b.read_state()


def test_walls_gravity_mixture_match_the_twin():
    """checks.rs walls_reversal's system (hard walls, gravity, binary mixture), vs the twin."""
    sys = rh.System.new(rh.box_for(3000, 0.4, 1.0), 0.005, rh.Pair.lj(2.5), [1.0, 3.0])
    sys = sys.with_pair(0, 1, rh.Pair(0.5, 1.2, 3.0)).with_hard_walls(0).with_hard_walls(1).with_gravity((0.0, -0.05))
    s0 = rh.fluid(sys, 3000, 1.0, 7)
    s0.species[::4] = 1
    gpu = rh.GpuEngine(sys, s0)
    gpu.step(400)
    assert gpu.read_state() == oracle_run(sys, s0, 400)
    gpu.reverse()
    gpu.step(400)
    gpu.reverse()
    assert gpu.read_state() == s0


# ---- all-integer pair forces (md::Arith::Int; md_core::table::IntTables) -----------------------

def golden_setup():
    """crates/md/src/checks.rs:156-164 (golden_int_setup), built with the product's md-core."""
    box = rh.box_for(3000, 0.6, 1.0)
    sys = rh.System.new(box, 0.005, rh.Pair.lj(2.5), [1.0, 2.0]).with_pair(0, 1, rh.Pair(0.7, 1.05, 2.5))
    st = rh.fluid(sys, 3000, 1.0, 2026)
    st.species[::4] = 1
    st.vel = rh.maxwell2(sys, st.species, 1.0, 2026)
    return sys, st


def test_integer_kicks_match_the_twin():
    sys, st = golden_setup()
    gpu = rh.GpuEngine(sys, st, rh.ForceMode.cell_list())
    gpu.set_arith(1)
    k, pe = gpu.kicks()
    want, pe_want = O.kicks2(oracle_system(sys), st.pos, st.species, prec=0)
    assert np.array_equal(k, want)
    assert abs(pe - pe_want) <= 1e-9 * abs(pe_want)


@pytest.mark.parametrize("verlet", [False, True])
def test_golden_integer_hash_on_the_gpu(verlet):
    """The reference's one stored trajectory value: GOLDEN_INT_HASH after 1000 all-integer
    steps of golden_int_setup() (crates/md/src/checks.rs:156-167, crates/md/tests/integer.rs:84-98),
    in cell-list mode and over Verlet lists (which must give the same trajectory)."""
    sys, st = golden_setup()
    mode = rh.verlet_with_skin(sys, st.n(), 0.3) if verlet else rh.ForceMode.cell_list()
    gpu = rh.GpuEngine(sys, st, mode)
    gpu.step(7)  # f32 steps first, then switch: the list is rebuilt in the integer path's form
    gpu.upload(st)
    gpu.set_arith(1)
    gpu.step(400)
    gpu.step(600)
    assert gpu.read_state().hash() == 0x3AAEFF680A2F0278
    if verlet:
        assert gpu.read_list_stats().rebuilds > 10
