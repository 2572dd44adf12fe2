"""The oracle's 2D engine with md_core::System's full boundaries (crates/md-ref/src/int_engine.rs,
crates/md-core/src/{system,walls}.rs), pinned by closed forms and invariants that a plausible
mistake breaks.  The reference keeps no stored vectors for these features.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

ONE = 1 << 32


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def walled_fluid(lx, ly, n, t=1.0, seed=3, masses=(1.0,)):
    """A 2D LJ fluid in an lx x ly sigma box with hard walls on both axes."""
    sys = O.system2([lx * ONE, ly * ONE], masses=masses)
    sys.walled[0] = sys.walled[1] = 1
    pos, vel, sp = O.fluid2(sys, n, t, seed)
    return sys, pos, vel, sp


def energy(sys, pos, vel, sp):
    ke, pe = O.energy2(sys, pos, vel, sp, 64)
    return ke + pe


def test_compartments_keep_particles_home_and_insulate():
    sys, pos, vel, sp = walled_fluid(64, 32, 900, t=1.5)
    sys.compartments[0] = 32 + 4  # 16 sigma compartments along x: four of them
    home = pos[:, 0] >> 36
    st = O.state2(sys)
    O.run2(sys, st, pos, vel, sp, 400)
    assert np.array_equal(pos[:, 0] >> 36, home)
    # compartment walls are exact mirrors: the run reverses bit for bit
    p1, v1 = pos.copy(), vel.copy()
    O.run2(sys, st, pos, vel, sp, 200)
    O.reverse_run2(sys, st, pos, vel, sp)
    O.run2(sys, st, pos, vel, sp, 200)
    O.reverse_run2(sys, st, pos, vel, sp)
    assert np.array_equal(pos, p1) and np.array_equal(vel, v1)


def test_insulating_compartments_remove_exactly_the_cross_pairs():
    """Kicks with insulating compartments = kicks of each compartment's particles on their own."""
    sys, pos, vel, sp = walled_fluid(64, 32, 900, t=1.5)
    sys.compartments[0] = 32 + 4
    st = O.state2(sys)
    k, _, _, _ = O.forces2(sys, st, pos, sp)
    home = pos[:, 0] >> 36
    for c in range(4):
        sel = np.where(home == c)[0]
        sub = O.system2([64 * ONE, 32 * ONE])
        sub.walled[0] = sub.walled[1] = 1
        kc, _, _, _ = O.forces2(sub, O.state2(sub), pos[sel], sp[sel])
        assert np.array_equal(k[sel], kc)
    sys.diathermal = 1
    kd, _, _, _ = O.forces2(sys, st, pos, sp)
    sub = O.system2([64 * ONE, 32 * ONE])
    sub.walled[0] = sub.walled[1] = 1
    kw, _, _, _ = O.forces2(sub, O.state2(sub), pos, sp)
    assert np.array_equal(kd, kw)  # diathermal walls do not change the pair forces


def test_membrane_blocks_one_species_and_pores_pass_it():
    sys, pos, vel, sp = walled_fluid(48, 48, 1200, t=1.5, masses=(1.0, 1.0))
    sp[::2] = 1
    vel[:] = O.maxwell2(sys, sp, 1.5, 4)
    sys.has_membrane = 1
    sys.membrane_axis = 0
    sys.membrane_blocks = 1 << 1
    sys.membrane_position = 24 * ONE
    # start with species 1 on the left, species 0 anywhere
    keep = ~((sp == 1) & (pos[:, 0] >= 24 * ONE))
    pos, vel, sp = pos[keep], vel[keep], sp[keep]
    st = O.state2(sys)
    O.run2(sys, st, pos, vel, sp, 600)
    assert (pos[sp == 1, 0] < 24 * ONE).all()           # blocked species stays on its side
    assert (pos[sp == 0, 0] >= 24 * ONE).any()           # the other crosses
    assert st.ledger.hits[4] + st.ledger.hits[5] > 0     # the membrane's faces were hit
    # pores every 8 sigma, 3 sigma wide, let the blocked species through
    st2 = O.state2(sys)
    sys.pore_shift = 32 + 3
    sys.pore_width = 3 * ONE
    p2, v2 = pos.copy(), vel.copy()
    O.run2(sys, st2, p2, v2, sp, 1500)
    assert (p2[sp == 1, 0] >= 24 * ONE).any()


def two_body(sys_kind, k, r0, d):
    """Two particles bonded by one bond at separation d (sigma, along a tilted axis), no pair
    forces: the kick on particle 0 against the closed form of the bond law."""
    sys = O.system2([64 * ONE, 64 * ONE], pair_=O.Pair(1.0, 1.0, 1e-3))
    bonds = np.array([[0, 1, 0]], np.uint32)
    sys.nbond_types = 1
    sys.bond_fene[0] = 1 if sys_kind == "fene" else 0
    sys.bond_k[0] = k
    sys.bond_r0[0] = r0
    sys.nbonds = 1
    sys.bonds = bonds.ctypes.data
    u = np.array([0.8, 0.6])
    pos = np.array([[20 * ONE, 20 * ONE], [20 * ONE + int(round(d * u[0] * ONE)), 20 * ONE + int(round(d * u[1] * ONE))]], np.int64)
    sp = np.zeros(2, np.uint32)
    kk, pe, _, _ = O.forces2(sys, O.state2(sys), pos, sp, prec=64)
    return kk, pe, (pos[1] - pos[0]) / ONE, bonds


@pytest.mark.parametrize("kind,k,r0,d", [("harmonic", 50.0, 1.0, 1.3), ("harmonic", 20.0, 0.0, 0.7), ("fene", 30.0, 1.5, 1.1)])
def test_bond_kicks_follow_the_bond_law(kind, k, r0, d):
    kk, pe, sep, _ = two_body(kind, k, r0, d)
    r = float(np.hypot(*sep))
    if kind == "harmonic":
        du = k * (r - r0)          # dU/dr of 1/2 k (r - r0)^2
        u = 0.5 * k * (r - r0) ** 2
    else:
        x = (r / r0) ** 2
        du = k * r / (1 - x)       # dU/dr of -1/2 k R0^2 ln(1 - (r/R0)^2)
        u = -0.5 * k * r0 * r0 * math.log(1 - x)
    c = 0.005 ** 2 * ONE
    want = du / r * c * sep        # particle 0 is pulled towards particle 1
    assert all(abs(kk[0, a] - want[a]) <= 1 + 1e-9 * abs(want[a]) for a in range(2))
    assert np.array_equal(kk[0], -kk[1])
    assert abs(pe - u) <= 1e-9 * max(1.0, u)


def test_bonded_chain_reverses_exactly_and_conserves_energy():
    sys, pos, vel, sp = walled_fluid(40, 40, 600, t=1.0)
    # chains of consecutive particles that start within 2 sigma of each other (lattice rows)
    chains = [[i, i + 1, 0] for i in range(0, 599) if (i % 10) != 9 and np.hypot(*((pos[i + 1] - pos[i]) / ONE)) < 2.0]
    bonds = np.array(chains, np.uint32)
    assert len(bonds) > 300
    sys.nbond_types = 1
    sys.bond_fene[0] = 0
    sys.bond_k[0] = 40.0
    sys.bond_r0[0] = 1.0
    sys.nbonds = len(bonds)
    sys.bonds = bonds.ctypes.data
    st = O.state2(sys)
    e0 = energy(sys, pos, vel, sp)
    p0, v0 = pos.copy(), vel.copy()
    O.run2(sys, st, pos, vel, sp, 300)
    assert abs(energy(sys, pos, vel, sp) - e0) < 2e-3 * abs(e0)
    O.reverse_run2(sys, st, pos, vel, sp)
    O.run2(sys, st, pos, vel, sp, 300)
    O.reverse_run2(sys, st, pos, vel, sp)
    assert np.array_equal(pos, p0) and np.array_equal(vel, v0)


def test_driven_piston_moves_at_its_speed_and_stops():
    sys, pos, vel, sp = walled_fluid(40, 40, 500, t=1.0)
    sys.has_piston = 1
    sys.piston_free = 0
    sys.piston_speed = -0.5
    sys.piston_start = 38.0
    sys.piston_min = 30.0
    sys.piston_max = 40.0
    st = O.state2(sys)
    v = st.piston_v
    assert v == int(round(-0.5 * 0.005 * ONE))
    w0 = st.piston_w
    O.run2(sys, st, pos, vel, sp, 100)
    assert st.piston_w == w0 + 100 * v                 # exact constant speed
    assert (pos[:, 1] < st.piston_w).all()
    O.run2(sys, st, pos, vel, sp, 4000)
    assert st.piston_w == 30 * ONE and st.piston_v == 0  # stopped at its lower limit
    assert (pos[:, 1] < 30 * ONE).all()


def test_free_piston_pressure_balance_and_work_bookkeeping():
    """A free piston under a constant load settles where the gas pushes back; the work done on
    it equals the integral of the gas force over its travel (piston_work)."""
    # the gas fills the 55 sigma below the piston's start; the box is 60 sigma tall
    low, pos, vel, sp = walled_fluid(30, 55, 300, t=1.0)
    sys = O.system2([30 * ONE, 60 * ONE])
    sys.walled[0] = sys.walled[1] = 1
    sys.has_piston = 1
    sys.piston_free = 1
    sys.piston_mass = 50.0
    sys.piston_force = 20.0
    sys.piston_start = 58.0
    sys.piston_min = 5.0
    sys.piston_max = 60.0
    st = O.state2(sys)
    ke0 = O.energy2(sys, pos, vel, sp, 64)[0]
    O.run2(sys, st, pos, vel, sp, 3000)
    h = st.piston_w / ONE
    assert 5.0 < h < 58.0 and st.piston_work < 0.0   # compressed: the gas did negative work
    assert (pos[:, 1] < st.piston_w).all()
    # the load did work F (58 - h) on the gas, which shows up as heat: the gas warmed
    assert O.energy2(sys, pos, vel, sp, 64)[0] > ke0


def test_fan_drives_its_slab_and_bath_layer_heats_it():
    sys, pos, vel, sp = walled_fluid(60, 30, 800, t=0.5)
    sys.walled[0] = 0  # periodic along x
    sys.has_bath = 1
    sys.bath_species = 0xFFFFFFFF
    sys.bath_t = 0.5
    sys.bath_rate = 5.0
    sys.bath_axis = 0
    sys.bath_region[0] = 0
    sys.bath_region[1] = 20 * ONE
    sys.bath_drift = 1.0
    st = O.state2(sys)
    O.run2(sys, st, pos, vel, sp, 3000)
    vx = vel[:, 0] / (0.005 * ONE)
    assert vx.mean() > 0.2                                  # the fan pushed the fluid along x
    assert st.ledger.hits[6] > 1000


def test_heat_bath_walls_book_the_energy_they_exchange():
    """With only heat-bath walls (no forces from walls), the change of the total energy equals
    the energy the walls booked, to the integrator's error."""
    sys, pos, vel, sp = walled_fluid(40, 40, 400, t=0.5)
    sys.wall_kind[1][0] = 2
    sys.wall_kind[1][1] = 2
    sys.wall_t[1][0] = 2.0
    sys.wall_t[1][1] = 2.0
    st = O.state2(sys)
    e0 = energy(sys, pos, vel, sp)
    O.run2(sys, st, pos, vel, sp, 2000)
    e1 = energy(sys, pos, vel, sp)
    booked = sum(st.ledger.energy[w] for w in range(4))
    assert st.ledger.hits[2] + st.ledger.hits[3] > 50
    assert booked > 0                                   # the hot walls heated the gas
    assert abs((e1 - e0) - booked) < 0.03 * abs(booked) + 1.0


def test_compartment_ledger_sums_to_the_totals():
    sys, pos, vel, sp = walled_fluid(64, 32, 900, t=1.5)
    sys.compartments[0] = 32 + 4
    sys.wall_kind[1][0] = 1  # a soft floor
    st = O.state2(sys)
    O.run2(sys, st, pos, vel, sp, 500)
    for w in range(4):
        assert st.ledger.hits[w] == sum(st.comp_hits[w][:])
        assert abs(st.ledger.impulse[w] - sum(st.comp_impulse[w][:])) <= 1e-9 * max(1.0, abs(st.ledger.impulse[w]))
    assert st.ledger.impulse[2] > 0  # the soft floor was pushed


def kinetic_per_particle(sys, vel, sp):
    m = np.array([sys.masses[s] for s in sp])
    v = vel.astype(np.float64) / ONE / sys.dt
    return float(np.sum(0.5 * m * (v ** 2).sum(axis=1)) / len(sp))


def test_thermostat_brings_a_cold_gas_to_its_temperature():
    """Andersen redraws at T: in 2D the mean kinetic energy per particle goes to kT (two degrees
    of freedom), for either mass (a wrong sd = T/m instead of sqrt(T/m) fails this)."""
    sys, pos, vel, sp = walled_fluid(30, 30, 200, t=0.2, masses=(1.0, 3.0))
    sp[1::2] = 1
    sys.has_bath = 1
    sys.bath_species = 0xFFFFFFFF
    sys.bath_t = 2.0
    sys.bath_rate = 10.0
    sys.bath_region[0] = -(1 << 63)
    sys.bath_region[1] = (1 << 63) - 1
    st = O.state2(sys)
    samples = []
    for _ in range(20):
        O.run2(sys, st, pos, vel, sp, 100)
        samples.append(kinetic_per_particle(sys, vel, sp))
    assert abs(np.mean(samples[5:]) - 2.0) < 0.15, np.mean(samples[5:])


def test_heat_bath_walls_bring_a_dilute_gas_to_their_temperature():
    """Every wall a heat bath at T: a dilute gas equilibrates to kT = T per particle."""
    sys, pos, vel, sp = walled_fluid(12, 12, 30, t=0.2)
    for a in range(2):
        for side in range(2):
            sys.wall_kind[a][side] = 2
            sys.wall_t[a][side] = 1.5
    st = O.state2(sys)
    samples = []
    for _ in range(60):
        O.run2(sys, st, pos, vel, sp, 200)
        samples.append(kinetic_per_particle(sys, vel, sp))
    assert abs(np.mean(samples[20:]) - 1.5) < 0.25, np.mean(samples[20:])


def test_free_piston_falls_freely_until_it_meets_the_gas():
    """With no particle within reach, a free piston under a load F follows the integer leapfrog
    exactly: v_k = k round(-F dt^2 / M 2^32), w_k = w_0 + v_1 + ... + v_k, and no work is done."""
    sys, pos, vel, sp = walled_fluid(20, 20, 40, t=0.5)
    pos[:, 1] = pos[:, 1] // 4  # the gas in the bottom quarter
    sys.box_len[1] = 60 * ONE
    sys.has_piston = 1
    sys.piston_free = 1
    sys.piston_mass = 7.0
    sys.piston_force = 3.0
    sys.piston_start = 59.0
    sys.piston_min = 1.0
    sys.piston_max = 60.0
    st = O.state2(sys)
    kick = int(round(-3.0 * 0.005 * 0.005 / 7.0 * ONE))
    O.run2(sys, st, pos, vel, sp, 200)
    assert st.piston_v == 200 * kick
    assert st.piston_w == int(round(59.0 * ONE)) + kick * 200 * 201 // 2
    assert st.piston_work == 0.0


# ---- md-ref's fast pair path (crates/md-ref/src/fast.rs) against the reference loop ----------

def _fast_cases():
    """The scenarios of crates/md-ref/tests/fast_path.rs:21-57: an LJ liquid; a WCA mixture of
    three masses and two sizes with hard walls on x, soft walls on y and gravity; bonded pairs
    in an LJ solvent."""
    liq = O.system2(O.box_for(2000, 0.7), masses=(1.0,))
    pos, vel, sp = O.fluid2(liq, 2000, 0.8, 1)
    yield "LJ liquid", liq, pos, vel, sp, None
    wca = lambda s: O.Pair(1.0, s, s * 2.0 ** (1.0 / 6.0))  # noqa: E731
    mix = O.system2(O.box_for(1500, 0.3, 1.5), masses=(1.0, 4.0, 30.0), pair_=wca(1.0))
    for a, b, sg in ((0, 2, 2.0), (1, 2, 2.0), (2, 2, 3.0)):
        mix.pairs[a * 4 + b] = wca(sg)
        mix.pairs[b * 4 + a] = wca(sg)
    mix.walled[0] = mix.walled[1] = 1
    mix.wall_kind[1][0] = mix.wall_kind[1][1] = 1  # soft floor and ceiling
    mix.gravity[1] = -0.01
    pos, vel, sp = O.fluid2(mix, 1500, 1.0, 2)
    sp[:] = [2 if i % 50 == 0 else i % 2 for i in range(1500)]
    big = pos[sp == 2]
    keep = [i for i in range(1500) if sp[i] == 2 or all(((b - pos[i]) / ONE) @ ((b - pos[i]) / ONE) > 4.0 for b in big)]
    pos, sp = np.ascontiguousarray(pos[keep]), np.ascontiguousarray(sp[keep])
    vel = O.maxwell2(mix, sp, 1.0, 2)
    yield "WCA mixture with walls and gravity", mix, pos, vel, sp, None
    chains = O.system2(O.box_for(1200, 0.5), masses=(1.0,))
    pos, vel, sp = O.fluid2(chains, 1200, 1.0, 3)
    bonds = np.array([[2 * k, 2 * k + 1, 0] for k in range(200)], np.uint32)
    chains.nbond_types = 1
    chains.bond_k[0] = 50.0
    chains.bond_r0[0] = 1.0
    chains.nbonds = len(bonds)
    chains.bonds = bonds.ctypes.data
    yield "bonded pairs in LJ", chains, pos, vel, sp, bonds


@pytest.mark.parametrize("prec", [32, 64])
def test_fast_pair_path_is_bit_identical_to_the_reference_loop(prec):
    """Trajectories of 300 steps and the kicks along them are identical bit for bit; the
    potential (a float sum in another order) agrees to 1e-4 (fast_path.rs:8-18)."""
    for name, sys, pos, vel, sp, keep_alive in _fast_cases():
        pa, va = pos.copy(), vel.copy()
        pb, vb = pos.copy(), vel.copy()
        with O.fast_path():
            ka, ea = O.kicks2(sys, pa, sp, prec)
            O.step2(sys, pa, va, sp, 300, prec)
            ka2, ea2 = O.kicks2(sys, pa, sp, prec)
        kb, eb = O.kicks2(sys, pb, sp, prec)
        O.step2(sys, pb, vb, sp, 300, prec)
        kb2, eb2 = O.kicks2(sys, pb, sp, prec)
        assert np.array_equal(ka, kb), name
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), f"{name}: trajectories differ"
        assert np.array_equal(ka2, kb2), name
        for x, y in ((ea, eb), (ea2, eb2)):
            assert abs(x - y) < 1e-4 * max(abs(y), 1.0), f"{name}: potential {x} against {y}"


def test_fast_path_is_exercised_and_not_the_reference_loop():
    """The switch really changes the code path: the kicks agree bit for bit, while the
    potential, summed pair by pair in the fast path's cell order instead of particle by
    particle, differs in its last bits."""
    sys = O.system2(O.box_for(2000, 0.7), masses=(1.0,))
    pos, vel, sp = O.fluid2(sys, 2000, 0.8, 1)
    slow, e_slow = O.kicks2(sys, pos, sp, 32)
    with O.fast_path():
        fast, e_fast = O.kicks2(sys, pos, sp, 32)
    assert np.array_equal(fast, slow)
    assert e_fast != e_slow and abs(e_fast - e_slow) < 1e-5 * abs(e_slow)
