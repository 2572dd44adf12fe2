// Device arithmetic shared by the single-system and batched engines: Morton keys, cell
// assignment, the exact minimum image, the reference's f32 separation, the per-pair integer kick
// (md-ref's operation order), soft walls and the wrap / mirror of the leapfrog.  Functions take
// any parameter struct with the fields they read (box, half, periodic, nc, cell_scale,
// soft_walls).
#pragma once

#include <cmath>

#include "prims.h"
#include "rh_common.h"

namespace rh {

constexpr float SOFT_RC_F = 1.122462048309373f;

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    uint32_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t spread2(uint32_t v) {
    uint32_t x = v & 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

template <int D>
__device__ __forceinline__ uint32_t morton(const uint32_t c[D]) {
    if constexpr (D == 3) return spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
    else return spread2(c[0]) | (spread2(c[1]) << 1);
}

template <int D, class PT>
__device__ __forceinline__ void cell_of(const PT& P, const int64_t x[D], uint32_t c[D]) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        double f = (double)x[a] * P.cell_scale[a];
        uint32_t ci = f > 0.0 ? (uint32_t)f : 0u;
        c[a] = ci < P.nc[a] - 1 ? ci : P.nc[a] - 1;
    }
}

// Minimum image (periodic axes) of b - a, exactly fixed::min_image (fixed.rs:31-40).
// Branch-free: the two corrections exclude each other, as the reference's if / else-if do.
template <int D, class PT>
__device__ __forceinline__ int64_t separation(const PT& P, int64_t a, int64_t b, int ax) {
    int64_t d = b - a;
    if ((P.periodic >> ax) & 1u) {
        int64_t down = d >= P.half[ax] ? -P.box[ax] : int64_t(0);
        int64_t up = d < -P.half[ax] ? P.box[ax] : int64_t(0);
        d += down + up;
    }
    return d;
}

// fixed::sep_f32 (fixed.rs:55-61): |d| >> 4 as f32 times 2^-28, with the sign of d; false if
// |d| >= 8 sigma.  Truncating d/16 toward zero and converting the signed value rounds
// symmetrically, so this equals the reference bit for bit (a tiny negative d gives +0 instead
// of -0, which no kick or energy can tell apart).
__device__ __forceinline__ bool sep_f(int64_t d, float& f) {
    bool ok = (uint64_t)(d + (int64_t(8) << 32)) < (uint64_t(16) << 32);
    int64_t q = (d + ((d >> 63) & 15)) >> 4;
    f = (float)(int32_t)q * 0x1p-28f;
    return ok;
}

// ---- The same operations on positions held in f64 ------------------------------------------
// The single-system engine keeps positions as f64 values that are integers (2^-32 sigma units)
// below 2^53 in magnitude, so every position, every difference of two positions and every
// position plus a step displacement is exact in f64, and the results below equal the i64
// versions bit for bit.  (The parameter struct adds boxf, halff and inv_box per axis.)

template <int D, class PT>
__device__ __forceinline__ void cell_of(const PT& P, const double x[D], uint32_t c[D]) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        double f = x[a] * P.cell_scale[a];
        uint32_t ci = f > 0.0 ? (uint32_t)f : 0u;
        c[a] = ci < P.nc[a] - 1 ? ci : P.nc[a] - 1;
    }
}

// Minimum image of b - a on periodic axes.  d - L rint(d / L) is exact (d and L rint(d / L)
// are integers below 2^53) and equals fixed::min_image's choice for every |d| other than L/2
// itself (where either image is as far as the box allows, beyond every cutoff the engine
// accepts, so no kick or energy depends on the choice).  Walled axes carry inv_box = 0, which
// makes the correction exactly zero (rint(0) = 0, d - L 0 = d) without a per-axis select.
template <int D, class PT>
__device__ __forceinline__ double separation(const PT& P, double a, double b, int ax) {
    const double d = b - a;
    return __builtin_fma(-P.boxf[ax], __builtin_rint(d * P.inv_box[ax]), d);
}

// fixed::sep_f32 on an f64 difference: trunc(d / 16) (exact: a power-of-two scale, then the
// integer part), rounded once to f32, times 2^-28; false if |d| >= 8 sigma.
__device__ __forceinline__ bool sep_f(double d, float& f) {
    const bool ok = d >= -34359738368.0 && d < 34359738368.0;
    f = (float)__builtin_trunc(d * 0.0625) * 0x1p-28f;
    return ok;
}

// The conversion alone, for list pairs whose separations are known to be in range.
__device__ __forceinline__ float sep_f_in_range(double d) { return (float)__builtin_trunc(d * 0.0625) * 0x1p-28f; }

// A half-list chunk's tag in the top four bits of its words (indices below 2^28): owner bits 0-3
// in x, 4-7 in y, 8-10 in z (bit 31 of z: in-block), entries - 1 in bits 28-29 of w and owner
// bit 11 in bit 30.
__device__ __forceinline__ uint4 tag_chunk(uint4 q, uint32_t own, uint32_t entries, bool in) {
    const uint32_t lo = 0x0fffffffu;
    return make_uint4((q.x & lo) | (own << 28), (q.y & lo) | ((own >> 4) << 28), (q.z & lo) | (((own >> 8) & 7u) << 28) | (in ? 1u << 31 : 0u),
                      (q.w & lo) | ((entries - 1u) << 28) | (((own >> 11) & 1u) << 30));
}

// An in-block chunk of the single-system half list: up to eight partners of one owner, as
// 12-bit offsets from the block's first slot, two to a word (bits 0-11 and 12-23); the top
// bytes hold the owner (12 bits, words x and y) and the entry count - 1 (word z).
__device__ __forceinline__ uint4 pack_inblock(const uint32_t e[8], uint32_t own, uint32_t entries) {
    const uint32_t m = 0xfffu;
    return make_uint4((e[0] & m) | ((e[1] & m) << 12) | ((own & 0xffu) << 24), (e[2] & m) | ((e[3] & m) << 12) | ((own >> 8) << 24),
                      (e[4] & m) | ((e[5] & m) << 12) | ((entries - 1u) << 24), (e[6] & m) | ((e[7] & m) << 12));
}

__device__ __forceinline__ void unpack_inblock(const uint4& q, uint32_t e[8], uint32_t& own, uint32_t& entries) {
    own = (q.x >> 24) | ((q.y >> 24) << 8);
    entries = ((q.z >> 24) & 7u) + 1u;
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 8; k++) e[k] = (w[k >> 1] >> (12 * (k & 1))) & 0xfffu;
}

// The indices of a tagged chunk, and its owner, entry count and in-block flag.
__device__ __forceinline__ uint4 untag_chunk(uint4 qt, uint32_t& own, uint32_t& entries, bool& in) {
    own = (qt.x >> 28) | ((qt.y >> 28) << 4) | (((qt.z >> 28) & 7u) << 8) | (((qt.w >> 30) & 1u) << 11);
    entries = ((qt.w >> 28) & 3u) + 1u;
    in = (qt.z >> 31) != 0u;
    return make_uint4(qt.x & 0x0fffffffu, qt.y & 0x0fffffffu, qt.z & 0x0fffffffu, qt.w & 0x0fffffffu);
}

// The same two steps on positions held in sixteenths (x / 16, exact for the integer positions).
// Every f64 quantity is its unscaled value times 2^-4 exactly, and so is every rounding: d * inv
// is the same product (2^-4 d times 2^4 inv), and the fma's single rounding scales with its
// result.  So separation16(a / 16, b / 16) = separation(a, b) / 16 bit for bit, and
// sep_f16(d / 16) = sep_f_in_range(d), one f64 multiply fewer per axis.
template <int D, class PT>
__device__ __forceinline__ double separation16(const PT& P, double a16, double b16, int ax) {
    const double d = b16 - a16;
    return __builtin_fma(-P.box16[ax], __builtin_rint(d * P.inv_box16[ax]), d);
}

__device__ __forceinline__ float sep_f16(double d16) { return (float)__builtin_trunc(d16) * 0x1p-28f; }

__device__ __forceinline__ bool sep_in_range(double d) { return d >= -34359738368.0 && d < 34359738368.0; }

template <int D, class PT>
__device__ __forceinline__ void soft_walls(const PT& P, const double x[D], float c, uint32_t k[D], double* wall_pe, float* wf = nullptr) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        for (int side = 0; side < 2; side++) {
            if (!((P.soft_walls >> (2 * a + side)) & 1u)) continue;
            double d = side == 0 ? x[a] : P.boxf[a] - x[a];
            float df;
            if (!sep_f(d, df)) continue;
            if (df < SOFT_RC_F) {
                float s2 = 1.0f / (df * df);
                float s6 = s2 * s2 * s2;
                float fw = 24.0f * s6 * (2.0f * s6 - 1.0f) / df;
                int32_t kw = cvt_sat_i32(__builtin_rintf(fw * c));
                k[a] += side == 0 ? (uint32_t)kw : (uint32_t)0 - (uint32_t)kw;
                if (wall_pe) *wall_pe += (double)(4.0f * s6 * (s6 - 1.0f) + 1.0f);
                if (wf) wf[2 * a + side] = fw;
            }
        }
    }
}

// Stochastic boundaries (md_core::three, three.rs:232-265): Threefry streams keyed by (particle
// id, clock) under the system seed.  The draws are shared by the engines that book exchanges in
// a ledger and those that keep none, so both move particles identically.
constexpr uint32_t STREAM_WALL = 0x57414c4cu;    // walls::STREAM_WALL
constexpr uint32_t STREAM_THERMO = 0x5448524du;  // walls::STREAM_THERMO
constexpr double TAU_D = 6.283185307179586;

// threefry::to_unit_f64 (threefry.rs:27-33): 53 random bits in [0, 1)
__device__ __forceinline__ double unit_f64(uint32_t r0, uint32_t r1) {
    const uint64_t bits = (((uint64_t)r0 << 21) ^ (uint64_t)(r1 >> 11)) & ((uint64_t(1) << 53) - 1u);
    return (double)bits * (1.0 / 9007199254740992.0);
}

// Rust `as i32` of round(v dt 2^32): a velocity (sigma/tau) as a displacement per step.
template <class PT>
__device__ __forceinline__ int32_t to_step(const PT& P, double v) {
    const double x = round(v * P.dt * 4294967296.0);
    if (x != x) return 0;
    if (x >= 2147483647.0) return INT32_MAX;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

template <class PT>
__device__ __forceinline__ double velocity(const PT& P, int32_t v) {
    return (double)v * (1.0 / 4294967296.0) / P.dt;
}

// md_core::three::andersen's draw: with probability rate dt, for a bath species that is not
// pinned, a fresh Maxwell velocity u (sigma/tau); false otherwise.
template <class PT>
__device__ __forceinline__ bool andersen_draw(const PT& P, uint32_t id, uint32_t clock, uint32_t s, double u[3]) {
    if (!((P.bath_species >> s) & 1u) || ((P.pinned >> s) & 1u)) return false;
    const uint32_t k0 = P.seed_lo, k1 = P.seed_hi ^ STREAM_THERMO;
    uint32_t r[2], r2[2], r3[2];
    threefry2x32(id, clock, k0, k1, r);
    if (unit_f64(r[0], 0u) >= P.bath_p) return false;
    threefry2x32(id, clock, k0, k1 ^ 1u, r2);
    threefry2x32(id, clock, k0, k1 ^ 2u, r3);
    const double sd = P.bath_sd[s];
    const double rad = sqrt(-2.0 * log(1.0 - unit_f64(r[1], 0u)));
    const double a = TAU_D * unit_f64(r2[0], 0u);
    const double rad3 = sqrt(-2.0 * log(1.0 - unit_f64(r3[0], 0u)));
    const double a3 = TAU_D * unit_f64(r3[1], 0u);
    u[0] = sd * rad * cos(a);
    u[1] = sd * rad * sin(a);
    u[2] = sd * rad3 * cos(a3);
    return true;
}

// The velocity a heat-bath wall w hands a particle of species s that crossed it (three.rs:232-243):
// the normal component pointing back in (side 0 is the low wall), the two tangential ones from
// one Box-Muller pair.
template <class PT>
__device__ __forceinline__ void thermal_wall_draw(const PT& P, uint32_t id, uint32_t clock, int w, int side, uint32_t s, double& vn, double& t0, double& t1) {
    const uint32_t k0 = P.seed_lo, k1 = P.seed_hi ^ STREAM_WALL;
    uint32_t r[2], r2[2];
    threefry2x32(id, clock, k0, k1, r);
    threefry2x32(id, clock, k0, k1 ^ 1u, r2);
    const double u1 = 1.0 - unit_f64(r[0], r[1]);
    const double u2 = 1.0 - unit_f64(r2[0], 0u), u3 = unit_f64(r2[1], 0u);
    const double sd = P.wall_sd[w][s];
    const double inward = side == 0 ? 1.0 : -1.0;
    vn = inward * sd * sqrt(-2.0 * log(u1));
    const double rad = sd * sqrt(-2.0 * log(u2));
    const double ang = TAU_D * u3;
    t0 = rad * cos(ang);
    t1 = rad * sin(ang);
}

// md_core::System's thermostat draw in 2D (the thermostat region of int_engine.rs's step): with
// probability rate dt a fresh Maxwell velocity u (sigma/tau) from one Box-Muller pair; false
// otherwise.  The caller checks the region, the species and pinning.
template <class PT>
__device__ __forceinline__ bool thermostat_draw2(const PT& P, uint32_t id, uint32_t clock, uint32_t s, double u[2]) {
    const uint32_t k0 = P.seed_lo, k1 = P.seed_hi ^ STREAM_THERMO;
    uint32_t r[2];
    threefry2x32(id, clock, k0, k1, r);
    if (!(unit_f64(r[0], 0u) < P.bath_p)) return false;
    uint32_t r2[2];
    threefry2x32(id, clock, k0, k1 ^ 1u, r2);
    const double u1 = 1.0 - unit_f64(r[1], 0u);
    const double u2 = unit_f64(r2[0], 0u);
    const double sd = P.bath_sd[s];
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = TAU_D * u2;
    u[0] = sd * rad * cos(ang);
    u[1] = sd * rad * sin(ang);
    return true;
}

// walls::thermal_velocity in 2D (walls.rs:91-104): the normal component from the wall-flux law,
// the tangential one a normal draw, under the system seed's wall stream.
template <class PT>
__device__ __forceinline__ void thermal_wall_draw2(const PT& P, uint32_t id, uint32_t clock, int w, int side, uint32_t s, double& vn, double& t) {
    const uint32_t k0 = P.seed_lo, k1 = P.seed_hi ^ STREAM_WALL;
    uint32_t r[2], r2[2];
    threefry2x32(id, clock, k0, k1, r);
    threefry2x32(id, clock, k0, k1 ^ 1u, r2);
    const double u1 = 1.0 - unit_f64(r[0], r[1]);
    const double u2 = 1.0 - unit_f64(r2[0], 0u), u3 = unit_f64(r2[1], 0u);
    const double sd = P.wall_sd[w][s];
    const double inward = side == 0 ? 1.0 : -1.0;
    vn = inward * sd * sqrt(-2.0 * log(u1));
    t = sd * sqrt(-2.0 * log(u2)) * cos(TAU_D * u3);
}

template <int D, class PT>
__device__ __forceinline__ void bound(const PT& P, double x[D], int32_t v[D]) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double l = P.boxf[a];
        if ((P.periodic >> a) & 1u) {
            if (x[a] < 0.0) x[a] += l;
            else if (x[a] >= l) x[a] -= l;
        } else if (x[a] < 0.0) {
            x[a] = -1.0 - x[a];
            v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
        } else if (x[a] >= l) {
            x[a] = 2.0 * l - 1.0 - x[a];
            v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
        }
    }
}

// The reciprocal and the division steps of the hardware's correctly rounded f32 division
// (v_rcp, one refinement of the reciprocal, then two correction steps), without its operand
// scaling and special-case fixup.  For operands where the hardware sequence scales nothing
// (normal divisor and quotient, exponent gap below 96, numerator above 2^-103) and the quotient
// is normal, this IS that sequence, so the result is the correctly rounded quotient.  The
// reciprocal is shared by the two divisions by r2 of a pair.
__device__ __forceinline__ float rcp_refined(float b) {
    float y0 = __builtin_amdgcn_rcpf(b);
    float e = __builtin_fmaf(-b, y0, 1.0f);
    return __builtin_fmaf(e, y0, y0);
}

__device__ __forceinline__ float div_refined(float a, float b, float y) {
    float q0 = a * y;
    float r0 = __builtin_fmaf(-b, q0, a);
    float q1 = __builtin_fmaf(r0, y, q0);
    float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y, q1);
}

// fr = 6 eps4 s6 (2 s6 - 1) / r2 with s2 = sigma2 / r2, in PairConsts::eval3's operation order
// (real.rs:145-156); 2*s6 is exact, so one fused op rounds (2*s6 - 1) once, as the separate ops
// do.  r2 >= r2_fast (checked for the whole wave) guarantees the shared-reciprocal division
// equals IEEE division (the host derives r2_fast per species pair); otherwise IEEE '/'.
__device__ __forceinline__ float pair_fr(float r2, float sigma2, float six_eps4, float r2_fast) {
    bool slow = !(r2 >= r2_fast);
    if (ballot64(slow) == 0ull) {
        float y = rcp_refined(r2);
        float s2 = div_refined(sigma2, r2, y);
        float s6 = s2 * s2 * s2;
        return div_refined(six_eps4 * s6 * __builtin_fmaf(2.0f, s6, -1.0f), r2, y);
    }
    float s2 = sigma2 / r2;
    float s6 = s2 * s2 * s2;
    return six_eps4 * s6 * __builtin_fmaf(2.0f, s6, -1.0f) / r2;
}

// One pair's kick on particle i: per-axis rint and the saturating conversion, summed with
// wrapping adds (int_engine3.rs:94-101); the factor 6*eps4 is the same f32 product md-ref forms.
template <int D>
__device__ __forceinline__ void pair_kick(const float f[D], float rc2, float sigma2, float six_eps4, float r2_fast, float c, uint32_t k[D]) {
    float r2 = f[0] * f[0] + f[1] * f[1];
    if constexpr (D == 3) r2 = r2 + f[2] * f[2];
    if (r2 < rc2) {
        float fr = pair_fr(r2, sigma2, six_eps4, r2_fast);
        float a = -fr * c;
#pragma unroll
        for (int ax = 0; ax < D; ax++) k[ax] += (uint32_t)cvt_sat_i32(__builtin_rintf(a * f[ax]));
    }
}

// The same pair with its energy u = eps4*s6*(s6-1) - shift.
template <int D>
__device__ __forceinline__ void pair_kick_pe(const float f[D], const PairTable* __restrict__ pt, int p, float c, uint32_t k[D], double& pe) {
    float r2 = f[0] * f[0] + f[1] * f[1];
    if constexpr (D == 3) r2 = r2 + f[2] * f[2];
    if (r2 < pt->rc2[p]) {
        float s2 = pt->sigma2[p] / r2;
        float s6 = s2 * s2 * s2;
        float fr = pt->six_eps4[p] * s6 * __builtin_fmaf(2.0f, s6, -1.0f) / r2;
        float a = -fr * c;
#pragma unroll
        for (int ax = 0; ax < D; ax++) k[ax] += (uint32_t)cvt_sat_i32(__builtin_rintf(a * f[ax]));
        float u = pt->eps4[p] * s6 * (s6 - 1.0f) - pt->shift[p];
        pe += (double)u;
    }
}

// Soft (WCA) walls on the centre's distance from each wall (int_engine3.rs:106-126).
// wf (optional, 2D entries): the force of each soft wall on the particle, for the ledger.
template <int D, class PT>
__device__ __forceinline__ void soft_walls(const PT& P, const int64_t x[D], float c, uint32_t k[D], double* wall_pe, float* wf = nullptr) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        for (int side = 0; side < 2; side++) {
            if (!((P.soft_walls >> (2 * a + side)) & 1u)) continue;
            int64_t d = side == 0 ? x[a] : P.box[a] - x[a];
            float df;
            if (!sep_f(d, df)) continue;
            if (df < SOFT_RC_F) {
                float s2 = 1.0f / (df * df);
                float s6 = s2 * s2 * s2;
                float fw = 24.0f * s6 * (2.0f * s6 - 1.0f) / df;
                int32_t kw = cvt_sat_i32(__builtin_rintf(fw * c));
                k[a] += side == 0 ? (uint32_t)kw : (uint32_t)0 - (uint32_t)kw;
                if (wall_pe) *wall_pe += (double)(4.0f * s6 * (s6 - 1.0f) + 1.0f);
                if (wf) wf[2 * a + side] = fw;
            }
        }
    }
}

// Wrap (periodic) or mirror about -1/2 and L-1/2 (walls), walls::mirror (walls.rs:27-35);
// a mirrored particle's velocity on that axis is reversed (hard and soft walls).
template <int D, class PT>
__device__ __forceinline__ void bound(const PT& P, int64_t x[D], int32_t v[D]) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        int64_t l = P.box[a];
        if ((P.periodic >> a) & 1u) {
            if (x[a] < 0) x[a] += l;
            else if (x[a] >= l) x[a] -= l;
        } else if (x[a] < 0) {
            x[a] = -1 - x[a];
            v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
        } else if (x[a] >= l) {
            x[a] = 2 * l - 1 - x[a];
            v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
        }
    }
}

template <int D, class PT>
__device__ __forceinline__ void stencil_cell(const PT& P, const uint32_t c[D], int o, uint32_t out[D]) {
    int off[3] = {o % 3 - 1, o / 3 % 3 - 1, o / 9 - 1};
#pragma unroll
    for (int a = 0; a < D; a++) {
        int v = (int)c[a] + off[a];
        if (v < 0) v += (int)P.nc[a];
        else if (v >= (int)P.nc[a]) v -= (int)P.nc[a];
        out[a] = (uint32_t)v;
    }
}

// ---- The ledger (md_core::walls::Ledger): wall and bath exchanges --------------------------
// Shared by the single-system and batched engines (the parameter struct adds mass, dt, the bath
// and wall draws' fields, thermal_walls and periodic).  Entries: 2 axis + side for the walls, 6
// for the bath; per entry energy, impulse and hits.

constexpr double LEDGER_ONE = 4294967296.0;      // ledger entries in 2^-32 units
constexpr int LEDGER_WORDS = 21;                 // 7 entries x (energy, impulse, hits)

// The ledger sums fixed-point values with 64-bit integer atomics, so its totals do not depend on
// the order of the additions.
__device__ __forceinline__ void ledger_add(unsigned long long* L, int w, int field, double value) {
    double q = rint(value * LEDGER_ONE);
    q = q != q ? 0.0 : fmin(fmax(q, -9.2e18), 9.2e18);  // saturate (a runaway wall force)
    atomicAdd(&L[3 * w + field], (unsigned long long)(long long)q);
}

// md_core::three::andersen: with probability rate dt a fresh Maxwell velocity, booked as the
// bath's (ledger entry 6) energy and hit.
template <int D, class PT>
__device__ void bath_redraw(const PT& P, unsigned long long* L, uint32_t id, uint32_t clock, uint32_t s, int32_t v[D]) {
    if constexpr (D == 3) {
        double u[3];
        if (!andersen_draw(P, id, clock, s, u)) return;
        double before = 0.0, after = 0.0;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const double w = velocity(P, v[b]);
            before += w * w;
        }
#pragma unroll
        for (int b = 0; b < 3; b++) v[b] = to_step(P, u[b]);
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const double w = velocity(P, v[b]);
            after += w * w;
        }
        ledger_add(L, 6, 0, 0.5 * P.mass[s] * (after - before));
        atomicAdd(&L[3 * 6 + 2], 1ull);
    }
}

// The 2D thermostat region with its fan: a particle of a coupled, unpinned species whose
// position along bath_axis is in [bath_lo, bath_hi) may take a fresh velocity plus the fan's
// drift along x, booked as the thermostat's (entry 6) energy and hit.  (Without a piston: the
// batched engine runs no piston.)
template <class PT>
__device__ void thermostat_redraw2(const PT& P, unsigned long long* L, uint32_t id, uint32_t clock, uint32_t s, const double xi[2], int32_t v[2]) {
    const double xb = P.bath_axis ? xi[1] : xi[0];
    if (((P.pinned >> s) & 1u) || !((P.bath_species >> s) & 1u) || !(xb >= P.bath_lo && xb < P.bath_hi)) return;
    double u[2];
    if (!thermostat_draw2(P, id, clock, s, u)) return;
    const double b0 = velocity(P, v[0]), b1 = velocity(P, v[1]);
    v[0] = to_step(P, u[0]);
    v[1] = to_step(P, u[1]);
    v[0] = (int32_t)((uint32_t)v[0] + (uint32_t)P.bath_drift);
    const double a0 = velocity(P, v[0]), a1 = velocity(P, v[1]);
    const double m = P.mass[s];
    ledger_add(L, 6, 0, 0.5 * m * (a0 * a0 + a1 * a1 - b0 * b0 - b1 * b1));
    atomicAdd(&L[3 * 6 + 2], 1ull);
}

// Wrap (periodic axes) or mirror about -1/2 and L - 1/2 (walls, walls.rs:27-35); a particle that
// crossed a wall leaves it with a reversed normal velocity (hard and soft) or a velocity drawn
// from the wall's bath (three.rs:232-243), and the exchange is booked in the ledger
// (int_engine3.rs:186-206).
template <int D, class PT>
__device__ void bound_booked(const PT& P, unsigned long long* L, double x[D], int32_t v[D], uint32_t id, uint32_t clock, uint32_t s) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double l = P.boxf[a];
        if ((P.periodic >> a) & 1u) {
            if (x[a] < 0.0) x[a] += l;
            else if (x[a] >= l) x[a] -= l;
            continue;
        }
        int side = -1;
        if (x[a] < 0.0) {
            x[a] = -1.0 - x[a];
            side = 0;
        } else if (x[a] >= l) {
            x[a] = 2.0 * l - 1.0 - x[a];
            side = 1;
        }
        if (side < 0) continue;
        const int w = 2 * a + side;
        double before[D], after[D];
#pragma unroll
        for (int b = 0; b < D; b++) before[b] = velocity(P, v[b]);
        if ((P.thermal_walls >> w) & 1u) {
            if constexpr (D == 3) {
                double vn, t0, t1;
                thermal_wall_draw(P, id, clock, w, side, s, vn, t0, t1);
                v[a] = to_step(P, vn);
                v[(a + 1) % 3] = to_step(P, t0);
                v[(a + 2) % 3] = to_step(P, t1);
            } else {
                double vn, t;
                thermal_wall_draw2(P, id, clock, w, side, s, vn, t);
                v[a] = to_step(P, vn);
                v[1 - a] = to_step(P, t);
            }
        } else {
            v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
        }
#pragma unroll
        for (int b = 0; b < D; b++) after[b] = velocity(P, v[b]);
        const double m = P.mass[s];
        double ka = after[0] * after[0], kb = before[0] * before[0];
#pragma unroll
        for (int b = 1; b < D; b++) {
            ka += after[b] * after[b];
            kb += before[b] * before[b];
        }
        ledger_add(L, w, 0, 0.5 * m * ka - 0.5 * m * kb);
        ledger_add(L, w, 1, fabs(m * (before[a] - after[a])));
        atomicAdd(&L[3 * w + 2], 1ull);
    }
}


constexpr int stencil_size(int d) { return d == 3 ? 27 : 9; }

// Bits u < t of a 32-bit mask (t may be out of [0, 32]).
__device__ __forceinline__ uint32_t bits_below(int t) {
    return t <= 0 ? 0u : (t >= 32 ? 0xffffffffu : ((1u << t) - 1u));
}

// 2 q + (bit of this lane in the wave mask m): one add with the mask as its carry-in.
__device__ __forceinline__ uint32_t shift_in(uint32_t q, uint64_t m) {
    uint32_t r;
    uint64_t carry_out;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(q), "s"(m));
    return r;
}

// The distance from both faces of a periodic axis, in sixteenths, beyond which a particle's
// listed partners need no periodic image: a list separation at step time is below the largest
// list radius plus the two particles' moves since the build (each under half the skin, or the
// list is rebuilt first), and 0.05 sigma covers the f32 tests.  (Host side.)
template <class PTab>
inline double face_margin16(const PTab& pt, float half_skin2) {
    float lr2 = 0.f;
    for (int k = 0; k < 16; k++) lr2 = lr2 > pt.list_r2[k] ? lr2 : pt.list_r2[k];
    return (std::sqrt((double)lr2) + 2.0 * std::sqrt((double)half_skin2) + 0.05) * 0x1p28;
}

// Entry k & 3 of a four-entry list chunk, without dynamic register indexing.
__device__ __forceinline__ void put4(uint4& q, uint32_t k, uint32_t j) {
    const uint32_t s = k & 3u;
    q.x = s == 0u ? j : q.x;
    q.y = s == 1u ? j : q.y;
    q.z = s == 2u ? j : q.z;
    q.w = s == 3u ? j : q.w;
}

// Which end of an in-block pair lists it: alternate by the parity of i + j so every particle
// owns about half of its in-block neighbours whatever its place in the block.
__device__ __forceinline__ bool owns(uint32_t i, uint32_t j) {
    uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
    return (((i ^ j) & 1u) ? lo : hi) == i;
}

// Kicks of one pair on both ends: adds k_i into ki and returns k_j in kj.  The partner's kick is
// rounded separately with its own factor (int_engine.rs:17-22): for equal factors it is the
// conversion of -rint(a f), the exact negation of this end's rounded value.
template <int D, bool MULTI>
__device__ __forceinline__ bool pair_kick2(const float f[D], float rc2, float sigma2, float six_eps4, float r2_fast, float ci, float cj, uint32_t ki[D], int32_t kj[D]) {
    float r2 = f[0] * f[0] + f[1] * f[1];
    if constexpr (D == 3) r2 = r2 + f[2] * f[2];
    const bool hit = r2 < rc2;  // kj is set only on a hit
    if (hit) {
        float fr = pair_fr(r2, sigma2, six_eps4, r2_fast);
        float a = -fr * ci;
        if constexpr (MULTI) {
            float aj = -fr * cj;
#pragma unroll
            for (int ax = 0; ax < D; ax++) {
                ki[ax] += (uint32_t)cvt_sat_i32(__builtin_rintf(a * f[ax]));
                kj[ax] = cvt_sat_i32(__builtin_rintf(aj * -f[ax]));
            }
        } else {
#pragma unroll
            for (int ax = 0; ax < D; ax++) {
                float r = __builtin_rintf(a * f[ax]);
                ki[ax] += (uint32_t)cvt_sat_i32(r);
                kj[ax] = cvt_sat_i32_neg(r);
            }
        }
    }
    return hit;
}

}  // namespace rh
