// md-core on the host (see core.h).  Compiled with -ffp-contract=off so each float operation
// rounds as the Rust originals do.
#include "core.h"

#include <cmath>
#include <limits>
#include <vector>

#include "engine.h"
#include "prims.h"

namespace rh::core {

int32_t sat_i32(double x) {
    if (std::isnan(x)) return 0;
    if (x >= 2147483647.0) return std::numeric_limits<int32_t>::max();
    if (x <= -2147483648.0) return std::numeric_limits<int32_t>::min();
    return (int32_t)x;
}

int64_t sat_i64(double x) {
    if (std::isnan(x)) return 0;
    if (x >= 9223372036854775807.0) return std::numeric_limits<int64_t>::max();
    if (x <= -9223372036854775808.0) return std::numeric_limits<int64_t>::min();
    return (int64_t)x;
}

uint32_t sat_u32(double x) {
    if (std::isnan(x) || x <= 0.0) return 0;
    if (x >= 4294967295.0) return std::numeric_limits<uint32_t>::max();
    return (uint32_t)x;
}

uint64_t sat_u64(double x) {
    if (std::isnan(x) || x <= 0.0) return 0;
    if (x >= 18446744073709551615.0) return std::numeric_limits<uint64_t>::max();
    return (uint64_t)x;
}

int64_t to_fixed(double x) { return sat_i64(std::round(x * ONE_F)); }

double powi(double x, int n) {
    // LLVM expands powi with a constant exponent by binary powering: the running product starts
    // at the lowest set bit's power and multiplies in each higher set bit's square.
    bool have = false;
    double acc = 1.0, sq = x;
    for (; n; n >>= 1) {
        if (n & 1) {
            acc = have ? acc * sq : sq;
            have = true;
        }
        if (n >> 1) sq = sq * sq;
    }
    return acc;
}

double pair_shift(const rh_pair& p) {
    if (p.eps == 0.0 || p.rc == 0.0) return 0.0;
    double s6 = powi(p.sigma / p.rc, 6);
    return 4.0 * p.eps * s6 * (s6 - 1.0);
}

double pair_f_over_r(const rh_pair& p, double r2) {
    if (r2 >= p.rc * p.rc) return 0.0;
    double s6 = powi(p.sigma * p.sigma / r2, 3);
    return 24.0 * p.eps * s6 * (2.0 * s6 - 1.0) / r2;
}

double pair_energy(const rh_pair& p, double r2) {
    if (r2 >= p.rc * p.rc) return 0.0;
    double s6 = powi(p.sigma * p.sigma / r2, 3);
    return 4.0 * p.eps * s6 * (s6 - 1.0) - pair_shift(p);
}

double to_unit_f64(uint32_t r0, uint32_t r1) {
    uint64_t bits = ((uint64_t)r0 << 21) ^ ((uint64_t)r1 >> 11);
    return (double)(bits & ((uint64_t(1) << 53) - 1)) * (1.0 / 9007199254740992.0);
}

double uniform(uint64_t seed, uint32_t stream, uint32_t i) {
    uint32_t r[2];
    threefry2x32(i, stream, (uint32_t)seed, (uint32_t)(seed >> 32), r);
    return to_unit_f64(r[0], r[1]);
}

// ---- the `libm` crate's routines (FreeBSD msun: e_log.c, k_cos.c, k_sin.c, e_rem_pio2.c) ----

namespace {

double bits(uint64_t b) {
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
uint64_t raw(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return b;
}
uint32_t high_word(double d) { return (uint32_t)(raw(d) >> 32); }

// Minimax coefficients, as IEEE bit patterns.
const uint64_t LOG_LG[7] = {0x3fe5555555555593ull, 0x3fd999999997fa04ull, 0x3fd2492494229359ull, 0x3fcc71c51d8e78afull,
                            0x3fc7466496cb03deull, 0x3fc39a09d078c69full, 0x3fc2f112df3e5244ull};
const uint64_t COS_C[6] = {0x3fa555555555554cull, 0xbf56c16c16c15177ull, 0x3efa01a019cb1590ull,
                           0xbe927e4f809c52adull, 0x3e21ee9ebdb4b1c4ull, 0xbda8fae9be8838d4ull};
const uint64_t SIN_S[6] = {0xbfc5555555555549ull, 0x3f8111111110f8a6ull, 0xbf2a01a019c161d5ull,
                           0x3ec71de357b1fe7dull, 0xbe5ae5e68a2b9cebull, 0x3de5d93a5acfd57cull};
// pi/2 split into three 33-bit pieces plus tails (pio2_1, pio2_1t, pio2_2, pio2_2t, pio2_3, pio2_3t)
const uint64_t PIO2[6] = {0x3ff921fb54400000ull, 0x3dd0b4611a626331ull, 0x3dd0b4611a600000ull,
                          0x3ba3198a2e037073ull, 0x3ba3198a2e000000ull, 0x397b839a252049c1ull};

double kernel_cos(double x, double y) {
    const double C1 = bits(COS_C[0]), C2 = bits(COS_C[1]), C3 = bits(COS_C[2]), C4 = bits(COS_C[3]), C5 = bits(COS_C[4]), C6 = bits(COS_C[5]);
    double z = x * x;
    double w = z * z;
    double r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6));
    double hz = 0.5 * z;
    double one_minus = 1.0 - hz;
    return one_minus + (((1.0 - one_minus) - hz) + (z * r - x * y));
}

double kernel_sin(double x, double y, bool tail) {
    const double S1 = bits(SIN_S[0]), S2 = bits(SIN_S[1]), S3 = bits(SIN_S[2]), S4 = bits(SIN_S[3]), S5 = bits(SIN_S[4]), S6 = bits(SIN_S[5]);
    double z = x * x;
    double w = z * z;
    double r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6);
    double v = z * x;
    if (!tail) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

// x - n*pi/2 for |x| up to 2^20*pi/2 (medium: Cody-Waite in up to three rounds).
int reduce_medium(double x, uint32_t ix, double& y0, double& y1) {
    const double inv_pio2 = bits(0x3fe45f306dc9c883ull);
    const double to_int = 1.5 / 2.220446049250313e-16;
    double tmp = x * inv_pio2 + to_int;
    double fn = tmp - to_int;
    int n = sat_i32(fn);
    double r = x - fn * bits(PIO2[0]);
    double w = fn * bits(PIO2[1]);
    y0 = r - w;
    int ex = (int)(ix >> 20);
    if (ex - (int)((raw(y0) >> 52) & 0x7ff) > 16) {
        double t = r;
        w = fn * bits(PIO2[2]);
        r = t - w;
        w = fn * bits(PIO2[3]) - ((t - r) - w);
        y0 = r - w;
        if (ex - (int)((raw(y0) >> 52) & 0x7ff) > 49) {
            t = r;
            w = fn * bits(PIO2[4]);
            r = t - w;
            w = fn * bits(PIO2[5]) - ((t - r) - w);
            y0 = r - w;
        }
    }
    y1 = (r - y0) - w;
    return n;
}

// One Cody-Waite step by k*pi/2 with the first split (small multiples, |x| <= 9pi/4).
int reduce_small(double x, int k, double& y0, double& y1) {
    const double p1 = bits(PIO2[0]), p1t = bits(PIO2[1]);
    double kk = (double)(k < 0 ? -k : k);
    if (k > 0) {
        double z = x - kk * p1;
        y0 = z - kk * p1t;
        y1 = (z - y0) - kk * p1t;
    } else {
        double z = x + kk * p1;
        y0 = z + kk * p1t;
        y1 = (z - y0) + kk * p1t;
    }
    return k;
}

// Returns false for |x| beyond the medium range (never produced by the engines' 2*pi*u).
bool rem_pio2(double x, int& n, double& y0, double& y1) {
    bool neg = (raw(x) >> 63) != 0;
    uint32_t ix = high_word(x) & 0x7fffffffu;
    int s = neg ? -1 : 1;
    if (ix <= 0x400f6a7au) {
        if ((ix & 0xfffffu) == 0x921fbu) n = reduce_medium(x, ix, y0, y1);
        else n = reduce_small(x, s * (ix <= 0x4002d97cu ? 1 : 2), y0, y1);
        return true;
    }
    if (ix <= 0x401c463bu) {
        if (ix <= 0x4015fdbcu) {
            if (ix == 0x4012d97cu) n = reduce_medium(x, ix, y0, y1);
            else n = reduce_small(x, s * 3, y0, y1);
        } else {
            if (ix == 0x401921fbu) n = reduce_medium(x, ix, y0, y1);
            else n = reduce_small(x, s * 4, y0, y1);
        }
        return true;
    }
    if (ix < 0x413921fbu) {
        n = reduce_medium(x, ix, y0, y1);
        return true;
    }
    return false;
}

}  // namespace

double libm_log(double x) {
    const double ln2_hi = bits(0x3fe62e42fee00000ull), ln2_lo = bits(0x3dea39ef35793c76ull);
    uint64_t u = raw(x);
    uint32_t hx = (uint32_t)(u >> 32);
    int k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((u << 1) == 0) return -1.0 / (x * x);
        if (hx >> 31) return (x - x) / 0.0;
        k = -54;
        x *= bits(0x4350000000000000ull);
        u = raw(x);
        hx = (uint32_t)(u >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (u << 32) == 0) {
        return 0.0;
    }
    // Reduce x to [sqrt(2)/2, sqrt(2)): x = 2^k (1 + f).
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    x = bits(((uint64_t)hx << 32) | (u & 0xffffffffull));
    double f = x - 1.0;
    double hfsq = 0.5 * f * f;
    double s = f / (2.0 + f);
    double z = s * s;
    double w = z * z;
    double t1 = w * (bits(LOG_LG[1]) + w * (bits(LOG_LG[3]) + w * bits(LOG_LG[5])));
    double t2 = z * (bits(LOG_LG[0]) + w * (bits(LOG_LG[2]) + w * (bits(LOG_LG[4]) + w * bits(LOG_LG[6]))));
    double r = t2 + t1;
    double dk = (double)k;
    return s * (hfsq + r) + dk * ln2_lo - hfsq + f + dk * ln2_hi;
}

double libm_cos(double x) {
    uint32_t ix = high_word(x) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e46a09eu && sat_i32(x) == 0) return 1.0;
        return kernel_cos(x, 0.0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    int n;
    double y0, y1;
    if (!rem_pio2(x, n, y0, y1)) return std::cos(x);
    switch (n & 3) {
    case 0: return kernel_cos(y0, y1);
    case 1: return -kernel_sin(y0, y1, true);
    case 2: return -kernel_cos(y0, y1);
    default: return kernel_sin(y0, y1, true);
    }
}

double libm_sin(double x) {
    uint32_t ix = high_word(x) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e500000u) return x;
        return kernel_sin(x, 0.0, false);
    }
    if (ix >= 0x7ff00000u) return x - x;
    int n;
    double y0, y1;
    if (!rem_pio2(x, n, y0, y1)) return std::sin(x);
    switch (n & 3) {
    case 0: return kernel_sin(y0, y1, true);
    case 1: return kernel_cos(y0, y1);
    case 2: return -kernel_sin(y0, y1, true);
    default: return -kernel_cos(y0, y1);
    }
}

double normal(uint64_t seed, uint32_t stream, uint32_t i) {
    const double tau = 6.283185307179586;
    double u1 = 1.0 - uniform(seed, stream, 2 * i);
    double u2 = uniform(seed, stream, 2 * i + 1);
    return std::sqrt(-2.0 * libm_log(u1)) * libm_cos(tau * u2);
}

double rc_max(const rh_pair* pairs, uint32_t ns) {
    double m = 0.0;
    for (uint32_t a = 0; a < ns; a++)
        for (uint32_t b = 0; b < ns; b++) m = std::fmax(m, pairs[a * 4 + b].rc);
    return m;
}

double kick_coeff(double dt, double mass, bool pinned) { return pinned ? 0.0 : dt * dt / mass * ONE_F; }
int32_t to_step_displacement(double dt, double v) { return sat_i32(std::round(v * dt * ONE_F)); }
int32_t gravity_kick(double dt, double g) { return sat_i32(std::round(g * dt * dt * ONE_F)); }

IntTables int_tables(const rh_pair* pairs, uint32_t ns, const double* masses, uint32_t pinned, double dt) {
    IntTables t;
    t.ns = ns;
    const double rcm = rc_max(pairs, ns);
    const double r2_scale = (double)(uint64_t(1) << TAB_R2_BITS), step = (double)(uint64_t(1) << TAB_STEP_BITS);
    t.len = (sat_u64(rcm * rcm * r2_scale) >> TAB_STEP_BITS) + 2;
    t.data.assign((size_t)ns * ns * t.len, 0);
    for (uint32_t a = 0; a < ns; a++) {
        const double c = kick_coeff(dt, masses[a], (pinned >> a) & 1u);
        for (uint32_t b = 0; b < ns; b++) {
            const rh_pair& p = pairs[a * 4 + b];
            t.rc2[a * ns + b] = sat_u32(p.rc * p.rc * r2_scale);
            rh_pair continued = p;
            continued.rc = INFINITY;  // the force continued past the cutoff, which applies separately
            for (uint64_t k = 0; k < t.len; k++) {
                const double r2 = std::fmax((double)(k * (uint64_t(1) << TAB_STEP_BITS)) / r2_scale, 0.36);
                double f = 0.0;
                if (r2 < p.rc * p.rc + 2.0 * step / r2_scale) f = pair_f_over_r(continued, r2) * c * (double)(1u << TAB_T_BITS);
                const double r = std::fmin(std::fmax(std::round(f), -2147483648.0), 2147483647.0);
                t.data[((size_t)a * ns + b) * t.len + k] = (int32_t)r;
            }
        }
    }
    return t;
}

}  // namespace rh::core

namespace rh {

// Bounds for pair_fr's fast path (all quantities f32, as the kernel forms them):
//   s2 = sigma2 / r2 needs no scaling when sigma2 is 0 or in [2^-60, 2^30] and r2 in
//   [2^-24, 64]: quotient and reciprocal normal, exponent gap below 96;
//   X = six_eps4 s6 (2 s6 - 1) has |X| <= 2 |six_eps4| max(s6, s6^2); with |X| < 2^64 and
//   r2 >= 2^-24 the gap stays below 96 and X / r2 below 2^88;
//   a nonzero X is at least |six_eps4| (sigma2/rc2)^3 2^-24 (2 s6 - 1 is a multiple of 2^-24
//   or exceeds 1/2 in magnitude), which must stay above 2^-100 (numerator and quotient normal).
float fast_division_floor(float rc2, float sigma2, float six_eps4) {
    const double inf = std::numeric_limits<double>::infinity();
    double s = sigma2, e = std::fabs((double)six_eps4), rc = rc2;
    if (!(rc > 0.0) || rc > 64.0) return (float)inf;
    if (s != 0.0 && (s < std::ldexp(1.0, -60) || s > std::ldexp(1.0, 30))) return (float)inf;
    if (e != 0.0 && s != 0.0 && e * std::pow(s / rc, 3.0) * std::ldexp(1.0, -24) < std::ldexp(1.0, -99)) return (float)inf;
    double floor2 = std::ldexp(1.0, -24);
    if (e != 0.0 && s != 0.0) {
        // 2 e s6max^2 < 2^63 (a factor two of margin) with s6max = (s / r2)^3
        double r2min = s * std::pow(2.0 * e / std::ldexp(1.0, 63), 1.0 / 6.0);
        floor2 = std::max(floor2, r2min * 1.01);
    }
    return (float)floor2;
}

}  // namespace rh

// ---------------------------------------------------------------------------------------------
// Initial conditions behind the C ABI.
// ---------------------------------------------------------------------------------------------

using namespace rh::core;

namespace {

constexpr uint32_t STREAM_VELOCITY = 1, STREAM_JITTER = 2;

template <int D>
void zero_sum(uint64_t n, int32_t* vel, const uint32_t* species, uint32_t pinned) {
    // Subtract the Euclidean quotient from every moving particle, and one more from the first
    // `remainder` of them, so each axis sums to exactly zero (three.rs:376-389, init.rs:76-89).
    std::vector<uint64_t> moving;
    moving.reserve(n);
    for (uint64_t i = 0; i < n; i++)
        if (!species || !((pinned >> species[i]) & 1u)) moving.push_back(i);
    int64_t m = (int64_t)moving.size();
    if (m == 0) return;
    for (int a = 0; a < D; a++) {
        int64_t sum = 0;
        for (uint64_t i : moving) sum += vel[D * i + a];
        int64_t q = sum / m, r = sum % m;
        if (r < 0) {
            r += m;
            q -= 1;
        }
        for (int64_t k = 0; k < m; k++) {
            uint64_t i = moving[k];
            vel[D * i + a] = (int32_t)((int64_t)vel[D * i + a] - (q + (k < r ? 1 : 0)));
        }
    }
}

template <int D>
uint64_t fnv_state(uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto eat = [&h](uint64_t v, int nbytes) {
        for (int b = 0; b < nbytes; b++) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= 0x100000001b3ull;
        }
    };
    for (uint64_t i = 0; i < n; i++) {
        if (D == 3) {
            for (int a = 0; a < 3; a++) {
                eat((uint64_t)pos[3 * i + a], 8);
                eat((uint32_t)vel[3 * i + a], 4);
            }
        } else {
            eat((uint64_t)pos[2 * i], 8);
            eat((uint64_t)pos[2 * i + 1], 8);
            eat((uint32_t)vel[2 * i], 4);
            eat((uint32_t)vel[2 * i + 1], 4);
        }
        eat(species[i], 4);
    }
    return h;
}

}  // namespace

extern "C" {

int rh_cube_for(uint64_t n, double density, int64_t out[3]) {
    int64_t l = to_fixed(std::cbrt((double)n / density));
    out[0] = out[1] = out[2] = l;
    return RH_OK;
}

int rh_fcc_cells(uint64_t n, const int64_t box_len[3], uint64_t c[3]) {
    double b[3] = {to_f64(box_len[0]), to_f64(box_len[1]), to_f64(box_len[2])};
    double a = std::cbrt(b[0] * b[1] * b[2] * 4.0 / (double)n);
    for (int k = 0; k < 3; k++) c[k] = std::max<uint64_t>(sat_u64(std::round(b[k] / a)), 1);
    while (4 * c[0] * c[1] * c[2] < n) {
        int widest = 0;
        for (int k = 1; k < 3; k++)
            if (b[k] / (double)c[k] > b[widest] / (double)c[widest]) widest = k;
        c[widest]++;
    }
    return RH_OK;
}

int rh_fcc(uint64_t n, const int64_t box_len[3], const uint64_t cells[3], double jitter, uint64_t seed, int64_t* pos) {
    static const double basis[4][3] = {{0.25, 0.25, 0.25}, {0.75, 0.75, 0.25}, {0.75, 0.25, 0.75}, {0.25, 0.75, 0.75}};
    double a[3];
    for (int x = 0; x < 3; x++) a[x] = to_f64(box_len[x]) / (double)cells[x];
    uint64_t sites = 4 * cells[0] * cells[1] * cells[2];
    if (sites < n) return RH_ERR_INVALID;
    for (uint64_t id = 0; id < n; id++) {
        uint64_t cell = id / 4, s = id % 4;
        double c[3] = {(double)(cell % cells[0]), (double)(cell / cells[0] % cells[1]), (double)(cell / cells[0] / cells[1])};
        for (int ax = 0; ax < 3; ax++) {
            double jit = jitter * (uniform(seed, STREAM_JITTER, (uint32_t)(3 * id + ax)) - 0.5);
            pos[3 * id + ax] = wrap(to_fixed((c[ax] + basis[s][ax]) * a[ax] + jit), box_len[ax]);
        }
    }
    return RH_OK;
}

int rh_maxwell3(const rh_system3* sys, uint64_t n, const uint32_t* species, double t, uint64_t seed, int32_t* vel) {
    for (uint64_t i = 0; i < n; i++) {
        uint32_t s = species[i];
        if (s >= sys->nspecies) return RH_ERR_INVALID;
        if ((sys->pinned >> s) & 1u) {
            vel[3 * i] = vel[3 * i + 1] = vel[3 * i + 2] = 0;
            continue;
        }
        double sd = std::sqrt(t / sys->masses[s]);
        for (int a = 0; a < 3; a++) vel[3 * i + a] = to_step_displacement(sys->dt, sd * normal(seed, STREAM_VELOCITY, (uint32_t)(3 * i + a)));
    }
    zero_sum<3>(n, vel, species, sys->pinned);
    return RH_OK;
}

int rh_set_temperature3(const rh_system3* sys, uint64_t n, const uint32_t* species, int32_t* vel, double t) {
    uint64_t moving = 0;
    double ke = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        if ((sys->pinned >> species[i]) & 1u) continue;
        moving++;
        double s = 0.0;
        for (int a = 0; a < 3; a++) s += powi((double)vel[3 * i + a] * UNIT / sys->dt, 2);
        ke += 0.5 * sys->masses[species[i]] * s;
    }
    double dof = 1.5 * std::fmax((double)moving - 1.0, 1.0);
    double now = ke / dof;
    if (now <= 0.0) return RH_OK;
    double f = std::sqrt(t / now);
    for (uint64_t i = 0; i < 3 * n; i++) vel[i] = sat_i32(std::round((double)vel[i] * f));
    zero_sum<3>(n, vel, species, sys->pinned);
    return RH_OK;
}

uint64_t rh_hash3(uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species) { return fnv_state<3>(n, pos, vel, species); }

int rh_box_for(uint64_t n, double density, double aspect, int64_t out[2]) {
    double ly = std::sqrt((double)n / density / aspect);
    out[0] = to_fixed(ly * aspect);
    out[1] = to_fixed(ly);
    return RH_OK;
}

int rh_lattice(uint64_t n, const int64_t box_len[2], double jitter, uint64_t seed, int64_t* pos) {
    double b0 = to_f64(box_len[0]), b1 = to_f64(box_len[1]);
    uint64_t nx = std::max<uint64_t>(sat_u64(std::ceil(std::sqrt((double)n * b0 / b1))), 1);
    uint64_t ny = (n + nx - 1) / nx;
    double a0 = b0 / (double)nx, a1 = b1 / (double)ny;
    for (uint64_t k = 0; k < n; k++) {
        double jx = jitter * (uniform(seed, STREAM_JITTER, (uint32_t)(2 * k)) - 0.5);
        double jy = jitter * (uniform(seed, STREAM_JITTER, (uint32_t)(2 * k + 1)) - 0.5);
        pos[2 * k] = wrap(to_fixed(((double)(k % nx) + 0.5) * a0 + jx), box_len[0]);
        pos[2 * k + 1] = wrap(to_fixed(((double)(k / nx) + 0.5) * a1 + jy), box_len[1]);
    }
    return RH_OK;
}

int rh_maxwell2(const rh_system2* sys, uint64_t n, const uint32_t* species, double t, uint64_t seed, int32_t* vel) {
    for (uint64_t i = 0; i < n; i++) {
        if (species[i] >= sys->nspecies) return RH_ERR_INVALID;
        double sd = std::sqrt(t / sys->masses[species[i]]);
        for (int a = 0; a < 2; a++) vel[2 * i + a] = to_step_displacement(sys->dt, sd * normal(seed, STREAM_VELOCITY, (uint32_t)(2 * i + a)));
    }
    zero_sum<2>(n, vel, nullptr, 0);
    return RH_OK;
}

int rh_fluid2(const rh_system2* sys, uint64_t n, double t, uint64_t seed, int64_t* pos, int32_t* vel, uint32_t* species) {
    int64_t inner[2];
    for (int a = 0; a < 2; a++) inner[a] = sys->walled[a] ? sys->box_len[a] - to_fixed(2.0) : sys->box_len[a];
    rh_lattice(n, inner, 0.1, seed, pos);
    for (uint64_t i = 0; i < n; i++) {
        for (int a = 0; a < 2; a++)
            if (sys->walled[a]) pos[2 * i + a] += to_fixed(1.0);
        species[i] = 0;
    }
    return rh_maxwell2(sys, n, species, t, seed, vel);
}

uint64_t rh_hash2(uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species) { return fnv_state<2>(n, pos, vel, species); }

int rh_verlet_mode3(const rh_system3* sys, uint64_t n, double skin, rh_mode* out) {
    double r = rc_max(sys->pairs, sys->nspecies) + skin;
    double vol = to_f64(sys->box_len[0]) * to_f64(sys->box_len[1]) * to_f64(sys->box_len[2]);
    double density = (double)n / vol;
    const double pi = 3.141592653589793;
    out->kind = 1;
    out->skin = skin;
    out->cap = sat_u32(std::ceil(4.0 / 3.0 * pi * r * r * r * std::fmax(2.0 * density, 1.2))) + 8;
    return RH_OK;
}

int rh_verlet_mode2(const rh_system2* sys, uint64_t n, double skin, rh_mode* out) {
    double r = rc_max(sys->pairs, sys->nspecies) + skin;
    double area = to_f64(sys->box_len[0]) * to_f64(sys->box_len[1]);
    double density = (double)n / area;
    const double pi = 3.141592653589793;
    out->kind = 1;
    out->skin = skin;
    out->cap = sat_u32(std::ceil(pi * r * r * std::fmax(2.0 * density, 1.2))) + 8;
    return RH_OK;
}

}  // extern "C"
