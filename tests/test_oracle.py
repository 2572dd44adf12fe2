"""The oracle (C restatement of md-ref / md-core) pinned against the reference's golden vectors,
known-answer tests and the identities its own unit tests assert.  CPU only."""
import json
import math
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

ONE = 1 << 32


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))

# crates/gpu-prims/src/threefry.rs:53-70 (Random123 kat_vectors)
KAT = [(r, tuple(int(x, 16) for x in c), tuple(int(x, 16) for x in k), tuple(int(x, 16) for x in o))
       for r, c, k, o in GOLDEN["threefry2x32_known_answers"]["vectors"]]


@pytest.mark.parametrize("rounds,ctr,key,want", KAT)
def test_threefry_known_answers(rounds, ctr, key, want):
    assert O.threefry(ctr, key, rounds) == want


def test_golden_all_integer_hash():
    """The reference's only stored trajectory value (crates/md/src/checks.rs:156-167): init
    (lattice, Threefry, libm Box-Muller), integer tables, the 2D leapfrog and the state hash."""
    sys, pos, vel, sp = O.golden_setup()
    O.step2(sys, pos, vel, sp, 1000, prec=0)
    assert O.hash2(pos, vel, sp) == O.GOLDEN_INT_HASH == int(GOLDEN["golden_int_hash"]["hash"], 16)


def _ulps(a: float, b: float) -> int:
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    return abs(int(ia) - int(ib))


def test_libm_port_tracks_the_c_library():
    """The FreeBSD-derived log/cos/sin agree with glibc to one ulp (they are different, both
    faithful algorithms); a wrong coefficient or reduction step would be off by many ulps."""
    rng = random.Random(5)
    worst, same, total = 0, 0, 0
    for _ in range(20000):
        u = rng.random()
        for f, g, x in ((O.lib().orc_log, math.log, 1.0 - u), (O.lib().orc_cos, math.cos, 2 * math.pi * u), (O.lib().orc_sin, math.sin, 2 * math.pi * u)):
            if x == 0.0:
                continue
            a, b = f(x), g(x)
            if abs(b) < 1e-300:
                continue
            d = _ulps(a, b)
            worst = max(worst, d)
            same += d == 0
            total += 1
    assert worst <= 1
    assert same / total > 0.95
    # special values
    assert O.lib().orc_log(1.0) == 0.0
    assert O.lib().orc_cos(0.0) == 1.0
    assert math.isinf(O.lib().orc_log(0.0))


def test_normal_moments():
    """threefry.rs:72-82"""
    xs = np.array([O.normal(42, 7, i) for i in range(50000)])
    assert abs(xs.mean()) < 0.02
    assert abs(xs.var() - 1.0) < 0.02


def test_fixed_identities():
    """crates/md-core/src/fixed.rs:74-99"""
    L = O.lib()
    l = 10 * ONE + 12345
    for a, b in [(0, l - 1), (l - 1, 0), (3, 3), (0, l // 2), (l // 2, 0), (1, l // 2 + 2)]:
        d = L.orc_min_image(a, b, l)
        assert -(l - l // 2) <= d < l // 2
        assert L.orc_wrap(a + d, l) == b
    assert L.orc_wrap(-1, l) == l - 1
    assert L.orc_wrap(l, l) == 0
    f1, f2 = np.zeros(1, np.float32), np.zeros(1, np.float32)
    for d in [1, 15, 16, 17, ONE, ONE + 7, 3 * ONE // 2, (8 << 32) - 1, 123_456_789]:
        assert L.orc_sep_f32(d, O.ptr(f1)) and L.orc_sep_f32(-d, O.ptr(f2))
        assert f2[0] == -f1[0]
    assert not L.orc_sep_f32(8 << 32, O.ptr(f1))
    assert L.orc_sep_f32(-(8 << 32), O.ptr(f1))
    L.orc_sep_f32(ONE + ONE // 4, O.ptr(f1))
    assert abs(f1[0] - 1.25) < 1e-7
    assert L.orc_to_fixed(1.5) == 3 * ONE // 2
    assert L.orc_to_fixed(-0.5 / ONE) == -1  # round half away from zero


def test_pair_closed_forms():
    """crates/md-core/src/pair.rs:63-83"""
    L = O.lib()
    wca = O.Pair(1.0, 1.0, 2.0 ** (1.0 / 6.0))
    rc2 = wca.rc * wca.rc
    import ctypes

    assert abs(L.orc_pair_energy(ctypes.byref(wca), rc2 * (1 - 1e-9))) < 1e-6
    assert abs(L.orc_pair_f_over_r(ctypes.byref(wca), rc2 * (1 - 1e-9))) < 1e-6
    assert abs(L.orc_pair_energy(ctypes.byref(wca), 1.0) - 1.0) < 1e-12
    lj = O.Pair(1.0, 1.0, 2.5)
    for r in [0.95, 1.1, 1.5, 2.2]:
        h = 1e-6
        du = (L.orc_pair_energy(ctypes.byref(lj), (r + h) ** 2) - L.orc_pair_energy(ctypes.byref(lj), (r - h) ** 2)) / (2 * h)
        assert abs(L.orc_pair_f_over_r(ctypes.byref(lj), r * r) * r + du) < 1e-6
    # shift: the unshifted energy at the cutoff
    s6 = (1.0 / 2.5) ** 6
    assert abs(L.orc_pair_shift(ctypes.byref(lj)) - 4.0 * s6 * (s6 - 1.0)) < 1e-15


def test_fcc_fills_the_box_without_overlaps():
    """crates/md-core/src/three.rs:410-430"""
    n = 500
    b = O.cube_for(n, 1.0)
    c = O.fcc_cells(n, b)
    assert list(c) == [5, 5, 5]
    p = O.fcc(n, b, c, 0.0, 1)
    l = int(b[0])
    d = p[:50, None, :] - p[None, :, :]
    d = (d + l // 2) % l - l // 2
    r = np.sqrt(((d.astype(np.float64) / ONE) ** 2).sum(-1))
    r[np.arange(50), np.arange(50)] = np.inf
    assert abs(r.min() - 4.0 ** (1 / 3) / math.sqrt(2)) < 1e-6


def test_fcc_cells_of_the_benchmark_sizes():
    """262,144 atoms at 0.8442 -> [41, 40, 40]; 4,194,304 -> [102, 102, 102] (the site
    counts that hold them)."""
    for n, want in [(262144, [41, 40, 40]), (4194304, [102, 102, 102])]:
        b = O.cube_for(n, 0.8442)
        assert list(O.fcc_cells(n, b)) == want


def test_maxwell_momentum_and_temperature():
    """crates/md-core/src/three.rs:432-446"""
    n = 4000
    box = O.cube_for(n, 0.5)
    sys = O.system3(box, pair_=O.Pair(1.0, 1.0, 2.0 ** (1 / 6)))
    sp = np.zeros(n, np.uint32)
    vel = O.maxwell3(sys, sp, 1.3, 2)
    assert (vel.astype(np.int64).sum(0) == 0).all()
    v = vel.astype(np.float64) / ONE / sys.dt
    t = 0.5 * (v ** 2).sum() / (1.5 * n)
    assert abs(t - 1.3) < 0.05
    O.set_temperature3(sys, vel, sp, 0.7)
    v = vel.astype(np.float64) / ONE / sys.dt
    t = 0.5 * (v ** 2).sum() / (1.5 * (n - 1))
    assert abs(t - 0.7) < 1e-3
    assert (vel.astype(np.int64).sum(0) == 0).all()


def test_int_tables_antisymmetric_and_accurate():
    """crates/md-core/src/table.rs:113-136"""
    import ctypes

    box = O.box_for(100, 0.5)
    sys = O.system2(box)
    tab = O.int_tables(sys)
    lj = O.Pair(1.0, 1.0, 2.5)
    c = 0.005 * 0.005 * ONE
    worst = 0.0
    for k in range(0, 2000, 3):
        r = 0.8 + 1.69 * k / 2000
        a = 0.3 + k * 0.001
        d = (O.lib().orc_to_fixed(r * math.cos(a)), O.lib().orc_to_fixed(r * math.sin(a)))
        ki = O.int_kick(sys, tab, 0, 0, d[0], d[1])
        kj = O.int_kick(sys, tab, 0, 0, -d[0], -d[1])
        assert ki == (-kj[0], -kj[1])
        f = O.lib().orc_pair_f_over_r(ctypes.byref(lj), r * r) * c
        exact = (-f * r * math.cos(a), -f * r * math.sin(a))
        for ax in range(2):
            worst = max(worst, abs(ki[ax] - exact[ax]) / (20.0 + 1e-5 * abs(exact[ax])))
    assert worst < 1.0


def _liquid3(n=600, density=0.8, t=1.0, walls=False):
    sys, pos, vel, sp = O.liquid3(n, density, t)
    if walls:
        sys.walled[2] = 1
        sys.wall_kind[2][0] = 0  # hard
        sys.wall_kind[2][1] = 0
        pos[:, 2] = np.clip(pos[:, 2], ONE // 2, int(sys.box_len[2]) - ONE // 2)
    return sys, pos, vel, sp


@pytest.mark.parametrize("prec", [32, 64])
def test_cell_list_equals_brute_force_3d(prec):
    """The stable cell list finds exactly the pairs an O(N^2) loop does: integer kicks agree
    bit for bit (integer sums do not depend on order)."""
    sys, pos, vel, sp = _liquid3(500, 0.8, 1.0)
    O.step3(sys, pos, vel, sp, 20, prec=32)
    k, _ = O.kicks3(sys, pos, sp, prec)
    kb = O.kicks3_brute(sys, pos, sp, prec)
    assert np.array_equal(k, kb)
    assert np.abs(k).max() > 0


def test_cell_list_equals_brute_force_3d_mixture_with_walls():
    sys, pos, vel, sp = _liquid3(1200, 0.7, 1.0, walls=True)
    sys.nspecies = 2
    sys.masses[1] = 3.0
    O.set_pair3(sys, 0, 1, O.Pair(0.5, 1.2, 3.0))
    sp[::3] = 1
    sys.wall_kind[2][1] = 1  # soft top wall
    O.step3(sys, pos, vel, sp, 10, prec=32)
    for prec in (32, 64):
        k, _ = O.kicks3(sys, pos, sp, prec)
        assert np.array_equal(k, O.kicks3_brute(sys, pos, sp, prec))


@pytest.mark.parametrize("prec", [0, 32, 64])
def test_cell_list_equals_brute_force_2d(prec):
    box = O.box_for(800, 0.7)
    sys = O.system2(box)
    pos, vel, sp = O.fluid2(sys, 800, 1.0, 3)
    O.step2(sys, pos, vel, sp, 30, prec=32)
    k, _ = O.kicks2(sys, pos, sp, prec)
    assert np.array_equal(k, O.kicks2_brute(sys, pos, sp, prec))


def test_oracle_reversal_is_exact_3d():
    """md-ref/tests/physics.rs:17-37 in 3D: n steps, reverse, n steps, reverse = start."""
    sys, pos, vel, sp = _liquid3(400, 0.8, 1.0)
    p0, v0 = pos.copy(), vel.copy()
    O.step3(sys, pos, vel, sp, 150)
    assert not np.array_equal(pos, p0)
    O.reverse3(sys, pos, vel, sp)
    O.step3(sys, pos, vel, sp, 150)
    O.reverse3(sys, pos, vel, sp)
    assert np.array_equal(pos, p0) and np.array_equal(vel, v0)


def test_oracle_momentum_and_permutation_invariance():
    """physics.rs:53-80: zero integer momentum stays zero; permuting storage permutes kicks."""
    sys, pos, vel, sp = _liquid3(400, 0.8, 1.0)
    O.step3(sys, pos, vel, sp, 50)
    assert (vel.astype(np.int64).sum(0) == 0).all()
    perm = np.random.default_rng(1).permutation(400)
    k, _ = O.kicks3(sys, pos, sp)
    kp, _ = O.kicks3(sys, np.ascontiguousarray(pos[perm]), np.ascontiguousarray(sp[perm]))
    assert np.array_equal(kp, k[perm])


def test_f32_kicks_track_f64():
    """physics.rs:129-146: rms relative difference of f32 and f64 kicks below 1e-5."""
    sys, pos, vel, sp = _liquid3(600, 0.8, 1.0)
    O.step3(sys, pos, vel, sp, 30)
    k32, pe32 = O.kicks3(sys, pos, sp, 32)
    k64, pe64 = O.kicks3(sys, pos, sp, 64)
    rel = math.sqrt(((k32 - k64).astype(np.float64) ** 2).sum() / (k64.astype(np.float64) ** 2).sum())
    assert rel < 1e-5
    assert abs(pe32 - pe64) / abs(pe64) < 1e-5


def test_planted_mistake_is_caught():
    """A check that can fail: perturbing one position by one unit changes the hash, and a
    reversed run no longer returns (gpu_engine.rs:91-105)."""
    sys, pos, vel, sp = _liquid3(600, 0.8, 1.0)
    p0, v0 = pos.copy(), vel.copy()
    h0 = O.hash3(pos, vel, sp)
    O.step3(sys, pos, vel, sp, 60)
    O.reverse3(sys, pos, vel, sp)
    pos[17, 1] ^= 1 << 20
    O.step3(sys, pos, vel, sp, 60)
    O.reverse3(sys, pos, vel, sp)
    assert O.hash3(pos, vel, sp) != h0
    assert not np.array_equal(pos, p0)


# ---- planted mistakes: each plausible error in the pair loop fails at least one pin ----------

def _lj_textbook(r, eps=1.0, sigma=1.0, rc=2.5):
    """U(r) = 4 eps ((s/r)^12 - (s/r)^6) - U(rc), the truncated and shifted LJ law."""
    def u(x):
        return 4.0 * eps * ((sigma / x) ** 12 - (sigma / x) ** 6)
    return u(r) - u(rc)


def _two_body_pin(prec=64):
    """Two atoms at a separation with three unequal components: each kick equals
    -dU/dr * (d/r) * dt^2/m * 2^32 from a central difference of the textbook law, and the
    potential equals U(r) (to f32's rounding at prec 32)."""
    box = [40 * ONE] * 3
    sys = O.system3(box)
    d = np.array([1.03, -0.41, 0.27])
    pos = np.array([[10 * ONE] * 3, [10 * ONE + int(round(x * ONE)) for x in d]], np.int64)
    sep = (pos[1] - pos[0]) / ONE
    r = float(np.sqrt((sep ** 2).sum()))
    h = 1e-6
    du = (_lj_textbook(r + h) - _lj_textbook(r - h)) / (2 * h)
    c = 0.005 ** 2 * ONE
    want = du / r * c * sep  # kick on atom 0: along +sep when attracted (du > 0)
    k, pe = O.kicks3(sys, pos, np.zeros(2, np.uint32), prec=prec)
    kicks_ok = all(abs(k[0, a] - want[a]) <= 1 + 1e-6 * abs(want[a]) for a in range(3)) and np.array_equal(k[0], -k[1])
    pe_ok = abs(pe - _lj_textbook(r)) <= (1e-9 if prec == 64 else 1e-6) * max(1.0, abs(_lj_textbook(r)))
    return kicks_ok and pe_ok


def _energy_pin(prec=64):
    """Total energy of a small liquid is conserved to 1e-3 over 200 steps of the engine at this
    precision (energies from the f64 evaluation)."""
    sys, pos, vel, sp = _liquid3(500, 0.8, 1.0)
    e0 = sum(O.energy3(sys, pos, vel, sp))
    O.step3(sys, pos, vel, sp, 200, prec)
    e1 = sum(O.energy3(sys, pos, vel, sp))
    return abs(e1 - e0) <= 1e-3 * abs(e0)


@pytest.mark.parametrize("prec", [32, 64])
def test_true_pair_law_passes_the_pins(prec):
    O.set_mutation(0)
    assert _two_body_pin(prec)
    assert _energy_pin(prec)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_planted_mistake_fails_a_pin(mutation, prec):
    """1 wrong sign, 2 dropped term, 3 wrong axis index, 4 dropped shift, 5 dropped r2 term:
    each fails a pin in the f32 engine (the GPU's twin) as in the f64 one."""
    O.set_mutation(mutation)
    try:
        caught = not _two_body_pin(prec) or not _energy_pin(prec)
    finally:
        O.set_mutation(0)
    assert caught, f"planted mistake {mutation} passed every pin at prec {prec}"


# ---- the state hash, pinned by FNV-1a's published vectors ---------------------------------

def _fnv1a64(data: bytes) -> int:
    """FNV-1a 64 (Fowler, Noll, Vo): offset basis 0xcbf29ce484222325, prime 0x100000001b3."""
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv1a_matches_its_published_vectors():
    assert _fnv1a64(b"") == 0xCBF29CE484222325
    assert _fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert _fnv1a64(b"foobar") == 0x85944171F73967E8


def test_state_hash_is_fnv1a_of_the_state_bytes():
    """State3::hash (three.rs:280-297): FNV-1a over, per particle, each axis's position (i64 LE)
    and velocity (i32 LE), then the species (u32 LE).  A 40-byte state spelled from a known string
    hashes to FNV-1a of that string; random states to FNV-1a of their bytes."""
    text = b"integer positions, integer velocities!\n\0"[:40]
    assert len(text) == 40
    pos = np.zeros((1, 3), np.int64)
    vel = np.zeros((1, 3), np.int32)
    off = 0
    for a in range(3):
        pos[0, a] = np.frombuffer(text[off:off + 8], "<i8")[0]
        vel[0, a] = np.frombuffer(text[off + 8:off + 12], "<i4")[0]
        off += 12
    sp = np.frombuffer(text[36:40], "<u4").copy()
    assert O.hash3(pos, vel, sp) == _fnv1a64(text)
    rng = np.random.default_rng(9)
    for n in (1, 2, 17):
        pos = rng.integers(-(1 << 62), 1 << 62, (n, 3), dtype=np.int64)
        vel = rng.integers(-(1 << 31), 1 << 31, (n, 3), dtype=np.int64).astype(np.int32)
        sp = rng.integers(0, 4, n, dtype=np.uint32)
        data = b"".join(b"".join(int(pos[i, a]).to_bytes(8, "little", signed=True) + int(vel[i, a]).to_bytes(4, "little", signed=True)
                                 for a in range(3)) + int(sp[i]).to_bytes(4, "little") for i in range(n))
        assert O.hash3(pos, vel, sp) == _fnv1a64(data)


# ---- the stochastic boundaries' draws (three.rs:232-265), pinned by their distributions ------

def test_heat_bath_wall_and_bath_draws_follow_their_distributions():
    """A heat-bath wall hands back the Maxwell flux distribution: the normal speed is Rayleigh
    with mean sqrt(pi T / 2 m) and second moment 2 T / m, the tangential components are
    Gaussian with variance T / m; the Andersen bath's three components are Gaussian with
    variance T / m.  200,000 draws, moments within five standard errors."""
    t, m, dt = 1.7, 2.5, 0.005
    n = 200_000
    d = O.draws3(0x1234_5678_9ABC, t, m, dt, 0, n, 41).astype(np.float64) / (dt * 2.0**32)
    sd2 = t / m
    bath, vn, tang = d[:, :3], d[:, 3], d[:, 4:]
    mean_vn = math.sqrt(math.pi * t / (2 * m))
    se = math.sqrt((2 - math.pi / 2) * sd2 / n)
    assert abs(vn.mean() - mean_vn) < 5 * se
    assert np.all(vn > 0)  # into the box from the low wall
    assert abs((vn ** 2).mean() - 2 * sd2) < 5 * 2 * sd2 / math.sqrt(n) * 1.5
    for col in (tang[:, 0], tang[:, 1], bath[:, 0], bath[:, 1], bath[:, 2]):
        assert abs(col.mean()) < 5 * math.sqrt(sd2 / n)
        assert abs(col.var() - sd2) < 5 * sd2 * math.sqrt(2.0 / n)
    # the two tangentials (one Box-Muller pair) are uncorrelated, and so are the bath's axes
    assert abs(np.corrcoef(tang[:, 0], tang[:, 1])[0, 1]) < 5 / math.sqrt(n)
    assert abs(np.corrcoef(bath[:, 0], bath[:, 2])[0, 1]) < 5 / math.sqrt(n)


def test_heat_bath_walls_book_the_energy_they_exchange_3d():
    """With heat-bath walls top and bottom (no wall forces), the change of a dilute gas's total
    energy equals the energy the walls booked in the ledger, to the integrator's error."""
    n = 300
    sys, pos, vel, sp = O.liquid3(n, 0.1, 0.5)
    sys.walled[2] = 1
    for side in (0, 1):
        sys.wall_kind[2][side] = 2
        sys.wall_t[2][side] = 2.0
    pos[:, 2] = np.clip(pos[:, 2], ONE // 2, int(sys.box_len[2]) - ONE // 2)
    e0 = sum(O.energy3(sys, pos, vel, sp))
    _, led = O.step3(sys, pos, vel, sp, 3000)
    e1 = sum(O.energy3(sys, pos, vel, sp))
    booked = led.energy[4] + led.energy[5]
    assert led.hits[4] + led.hits[5] > 50 and led.hits[0] == led.hits[1] == 0
    assert booked > 0  # the hot walls heated the gas
    assert abs((e1 - e0) - booked) < 0.03 * abs(booked) + 1.0
