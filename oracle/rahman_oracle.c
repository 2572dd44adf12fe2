/*
 * rahman_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * CPU restatement of md-ref (crates/md-ref) and the md-core / gpu-prims pieces it uses, written
 * from the reference's Rust sources.  Built with -O2 -ffp-contract=off so every float operation
 * rounds exactly as the Rust code's does.  See rahman_oracle.h for what pins each part.
 */
#include "rahman_oracle.h"

/* Planted mistakes, for the tests that show the parity pins catch them (0 = none):
 * 1 wrong sign of the pair force, 2 dropped "-1" of (2 s6 - 1), 3 wrong axis index in the kick,
 * 4 dropped energy shift, 5 dropped z term of r2.  Applies to the 3D pair loop only. */
int orc_mutation = 0;
void orc_set_mutation(int m) { orc_mutation = m; }

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ONE_F ((double)(1ULL << 32))
#define UNIT (1.0 / 4294967296.0)
#define SEP_LIMIT (8LL << 32)
#define SOFT_RC 1.122462048309373
#define TAU 6.283185307179586
#define STREAM_VELOCITY 1u
#define STREAM_JITTER 2u
#define STREAM_WALL 0x57414c4cu
#define STREAM_THERMO 0x5448524du

/* ------------------------------------------------------------------------------------------ */
/* Rust numeric conversions                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* `x as i32` for floats: truncation, saturating, NaN -> 0. */
static int32_t sat_i32(double x) {
    if (x != x) return 0;
    if (x >= 2147483647.0) return INT32_MAX;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

static int64_t sat_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}

static uint32_t sat_u32(double x) {
    if (x != x || x <= 0.0) return 0;
    if (x >= 4294967295.0) return UINT32_MAX;
    return (uint32_t)x;
}

static uint64_t sat_u64(double x) {
    if (x != x || x <= 0.0) return 0;
    if (x >= 18446744073709551615.0) return UINT64_MAX;
    return (uint64_t)x;
}

/* powi by binary powering, as LLVM expands constant exponents
 * (res starts as the first set bit's square, then res *= square for each later set bit). */
static double powi_d(double a, int b) {
    double res = 0.0;
    int have = 0;
    double sq = a;
    while (b) {
        if (b & 1) {
            res = have ? res * sq : sq;
            have = 1;
        }
        b >>= 1;
        if (b) sq = sq * sq;
    }
    return have ? res : 1.0;
}

/* ------------------------------------------------------------------------------------------ */
/* Threefry-2x32 (crates/gpu-prims/src/threefry.rs:3-46)                                      */
/* ------------------------------------------------------------------------------------------ */

static const uint32_t TF_ROT[8] = {13, 15, 26, 6, 17, 29, 16, 24};

static inline uint32_t rotl32(uint32_t x, uint32_t r) { return (x << r) | (x >> (32 - r)); }

static void threefry(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t rounds, uint32_t out[2]) {
    uint32_t ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
    uint32_t x0 = c0 + ks[0], x1 = c1 + ks[1];
    for (uint32_t r = 0; r < rounds; r++) {
        x0 += x1;
        x1 = rotl32(x1, TF_ROT[r % 8]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            uint32_t s = (r >> 2) + 1;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + s;
        }
    }
    out[0] = x0;
    out[1] = x1;
}

void orc_threefry2x32_rounds(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t rounds, uint32_t out[2]) {
    threefry(c0, c1, k0, k1, rounds, out);
}

/* threefry.rs:29-32 */
double orc_to_unit_f64(uint32_t r0, uint32_t r1) {
    uint64_t bits = ((uint64_t)r0 << 21) ^ ((uint64_t)r1 >> 11);
    return (double)(bits & ((1ULL << 53) - 1)) * (1.0 / (double)(1ULL << 53));
}

/* threefry.rs:35-37 */
double orc_uniform(uint64_t seed, uint32_t stream, uint32_t i) {
    uint32_t r[2];
    threefry(i, stream, (uint32_t)seed, (uint32_t)(seed >> 32), 20, r);
    return orc_to_unit_f64(r[0], r[1]);
}

/* ------------------------------------------------------------------------------------------ */
/* The `libm` crate's log / cos / sin (Rust ports of the FreeBSD msun / musl routines:         */
/* e_log.c, s_cos.c, s_sin.c, k_cos.c, k_sin.c, e_rem_pio2.c).  Written from those public      */
/* algorithms; constants given by their IEEE bit patterns.                                    */
/* ------------------------------------------------------------------------------------------ */

static inline double from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

double orc_log(double x) {
    const double ln2_hi = from_bits(0x3fe62e42fee00000ULL);
    const double ln2_lo = from_bits(0x3dea39ef35793c76ULL);
    const double lg1 = from_bits(0x3fe5555555555593ULL);
    const double lg2 = from_bits(0x3fd999999997fa04ULL);
    const double lg3 = from_bits(0x3fd2492494229359ULL);
    const double lg4 = from_bits(0x3fcc71c51d8e78afULL);
    const double lg5 = from_bits(0x3fc7466496cb03deULL);
    const double lg6 = from_bits(0x3fc39a09d078c69fULL);
    const double lg7 = from_bits(0x3fc2f112df3e5244ULL);
    uint64_t ui = to_bits(x);
    uint32_t hx = (uint32_t)(ui >> 32);
    int k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((ui << 1) == 0) return -1.0 / (x * x);
        if (hx >> 31) return (x - x) / 0.0;
        k -= 54;
        x *= from_bits(0x4350000000000000ULL);
        ui = to_bits(x);
        hx = (uint32_t)(ui >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (ui << 32) == 0) {
        return 0.0;
    }
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    ui = ((uint64_t)hx << 32) | (ui & 0xffffffffULL);
    x = from_bits(ui);
    double f = x - 1.0;
    double hfsq = 0.5 * f * f;
    double s = f / (2.0 + f);
    double z = s * s;
    double w = z * z;
    double t1 = w * (lg2 + w * (lg4 + w * lg6));
    double t2 = z * (lg1 + w * (lg3 + w * (lg5 + w * lg7)));
    double r = t2 + t1;
    double dk = (double)k;
    return s * (hfsq + r) + dk * ln2_lo - hfsq + f + dk * ln2_hi;
}

static double k_cos(double x, double y) {
    const double c1 = from_bits(0x3fa555555555554cULL);
    const double c2 = from_bits(0xbf56c16c16c15177ULL);
    const double c3 = from_bits(0x3efa01a019cb1590ULL);
    const double c4 = from_bits(0xbe927e4f809c52adULL);
    const double c5 = from_bits(0x3e21ee9ebdb4b1c4ULL);
    const double c6 = from_bits(0xbda8fae9be8838d4ULL);
    double z = x * x;
    double w = z * z;
    double r = z * (c1 + z * (c2 + z * c3)) + w * w * (c4 + z * (c5 + z * c6));
    double hz = 0.5 * z;
    w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * r - x * y));
}

static double k_sin(double x, double y, int iy) {
    const double s1 = from_bits(0xbfc5555555555549ULL);
    const double s2 = from_bits(0x3f8111111110f8a6ULL);
    const double s3 = from_bits(0xbf2a01a019c161d5ULL);
    const double s4 = from_bits(0x3ec71de357b1fe7dULL);
    const double s5 = from_bits(0xbe5ae5e68a2b9cebULL);
    const double s6 = from_bits(0x3de5d93a5acfd57cULL);
    double z = x * x;
    double w = z * z;
    double r = s2 + z * (s3 + z * s4) + z * w * (s5 + z * s6);
    double v = z * x;
    if (iy == 0) return x + v * (s1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * s1);
}

static const uint64_t PIO2_1_B = 0x3ff921fb54400000ULL;
static const uint64_t PIO2_1T_B = 0x3dd0b4611a626331ULL;
static const uint64_t PIO2_2_B = 0x3dd0b4611a600000ULL;
static const uint64_t PIO2_2T_B = 0x3ba3198a2e037073ULL;
static const uint64_t PIO2_3_B = 0x3ba3198a2e000000ULL;
static const uint64_t PIO2_3T_B = 0x397b839a252049c1ULL;

static int rem_pio2_medium(double x, uint32_t ix, double *y0o, double *y1o) {
    const double invpio2 = from_bits(0x3fe45f306dc9c883ULL);
    const double toint = 1.5 / 2.220446049250313e-16;
    const double pio2_1 = from_bits(PIO2_1_B), pio2_1t = from_bits(PIO2_1T_B);
    const double pio2_2 = from_bits(PIO2_2_B), pio2_2t = from_bits(PIO2_2T_B);
    const double pio2_3 = from_bits(PIO2_3_B), pio2_3t = from_bits(PIO2_3T_B);
    double tmp = x * invpio2 + toint;
    double fn = tmp - toint;
    int n = (int)sat_i32(fn);
    double r = x - fn * pio2_1;
    double w = fn * pio2_1t;
    double y0 = r - w;
    int ey = (int)((to_bits(y0) >> 52) & 0x7ff);
    int ex = (int)(ix >> 20);
    if (ex - ey > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y0 = r - w;
        ey = (int)((to_bits(y0) >> 52) & 0x7ff);
        if (ex - ey > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y0 = r - w;
        }
    }
    *y0o = y0;
    *y1o = (r - y0) - w;
    return n;
}

/* Returns n with x = n*pi/2 + y0 + y1.  Arguments beyond 2^20*pi/2 are never produced by the
 * engines (they pass 2*pi*u with u in [0,1)); those fall back to the C library. */
static int rem_pio2(double x, double *y0, double *y1, int *ok) {
    const double pio2_1 = from_bits(PIO2_1_B), pio2_1t = from_bits(PIO2_1T_B);
    int sign = (int)(to_bits(x) >> 63);
    uint32_t ix = (uint32_t)(to_bits(x) >> 32) & 0x7fffffffu;
    double z;
    *ok = 1;
    if (ix <= 0x400f6a7au) {
        if ((ix & 0xfffffu) == 0x921fbu) return rem_pio2_medium(x, ix, y0, y1);
        if (ix <= 0x4002d97cu) {
            if (!sign) {
                z = x - pio2_1;
                *y0 = z - pio2_1t;
                *y1 = (z - *y0) - pio2_1t;
                return 1;
            }
            z = x + pio2_1;
            *y0 = z + pio2_1t;
            *y1 = (z - *y0) + pio2_1t;
            return -1;
        }
        if (!sign) {
            z = x - 2.0 * pio2_1;
            *y0 = z - 2.0 * pio2_1t;
            *y1 = (z - *y0) - 2.0 * pio2_1t;
            return 2;
        }
        z = x + 2.0 * pio2_1;
        *y0 = z + 2.0 * pio2_1t;
        *y1 = (z - *y0) + 2.0 * pio2_1t;
        return -2;
    }
    if (ix <= 0x401c463bu) {
        if (ix <= 0x4015fdbcu) {
            if (ix == 0x4012d97cu) return rem_pio2_medium(x, ix, y0, y1);
            if (!sign) {
                z = x - 3.0 * pio2_1;
                *y0 = z - 3.0 * pio2_1t;
                *y1 = (z - *y0) - 3.0 * pio2_1t;
                return 3;
            }
            z = x + 3.0 * pio2_1;
            *y0 = z + 3.0 * pio2_1t;
            *y1 = (z - *y0) + 3.0 * pio2_1t;
            return -3;
        }
        if (ix == 0x401921fbu) return rem_pio2_medium(x, ix, y0, y1);
        if (!sign) {
            z = x - 4.0 * pio2_1;
            *y0 = z - 4.0 * pio2_1t;
            *y1 = (z - *y0) - 4.0 * pio2_1t;
            return 4;
        }
        z = x + 4.0 * pio2_1;
        *y0 = z + 4.0 * pio2_1t;
        *y1 = (z - *y0) + 4.0 * pio2_1t;
        return -4;
    }
    if (ix < 0x413921fbu) return rem_pio2_medium(x, ix, y0, y1);
    *ok = 0;
    return 0;
}

double orc_cos(double x) {
    uint32_t ix = (uint32_t)(to_bits(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e46a09eu) {
            if (sat_i32(x) == 0) return 1.0;
        }
        return k_cos(x, 0.0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    double y0, y1;
    int ok;
    int n = rem_pio2(x, &y0, &y1, &ok);
    if (!ok) return cos(x);
    switch (n & 3) {
    case 0: return k_cos(y0, y1);
    case 1: return -k_sin(y0, y1, 1);
    case 2: return -k_cos(y0, y1);
    default: return k_sin(y0, y1, 1);
    }
}

double orc_sin(double x) {
    uint32_t ix = (uint32_t)(to_bits(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e500000u) return x;
        return k_sin(x, 0.0, 0);
    }
    if (ix >= 0x7ff00000u) return x - x;
    double y0, y1;
    int ok;
    int n = rem_pio2(x, &y0, &y1, &ok);
    if (!ok) return sin(x);
    switch (n & 3) {
    case 0: return k_sin(y0, y1, 1);
    case 1: return k_cos(y0, y1);
    case 2: return -k_sin(y0, y1, 1);
    default: return -k_cos(y0, y1);
    }
}

/* threefry.rs:42-46: Box-Muller through the `libm` crate (sqrt is correctly rounded anyway). */
double orc_normal(uint64_t seed, uint32_t stream, uint32_t i) {
    double u1 = 1.0 - orc_uniform(seed, stream, 2 * i);
    double u2 = orc_uniform(seed, stream, 2 * i + 1);
    return sqrt(-2.0 * orc_log(u1)) * orc_cos(TAU * u2);
}

/* ------------------------------------------------------------------------------------------ */
/* md-core fixed point (crates/md-core/src/fixed.rs:21-72)                                     */
/* ------------------------------------------------------------------------------------------ */

int64_t orc_to_fixed(double x) { return sat_i64(round(x * ONE_F)); }
double orc_to_f64(int64_t x) { return (double)x * UNIT; }

int64_t orc_min_image(int64_t a, int64_t b, int64_t l) {
    int64_t d = b - a;
    int64_t half = l / 2;
    if (d >= half) d -= l;
    else if (d < -half) d += l;
    return d;
}

int64_t orc_wrap(int64_t x, int64_t l) {
    if (x < 0) return x + l;
    if (x >= l) return x - l;
    return x;
}

int orc_sep_f32(int64_t d, float *out) {
    if (!(d >= -SEP_LIMIT && d < SEP_LIMIT)) return 0;
    uint64_t m = d < 0 ? (uint64_t)0 - (uint64_t)d : (uint64_t)d;
    float mag = (float)(uint32_t)(m >> 4) * (1.0f / 268435456.0f);
    *out = d < 0 ? -mag : mag;
    return 1;
}

static int sep_f64(int64_t d, double *out) {
    if (!(d >= -SEP_LIMIT && d < SEP_LIMIT)) return 0;
    *out = orc_to_f64(d);
    return 1;
}

/* ------------------------------------------------------------------------------------------ */
/* md-core Pair (crates/md-core/src/pair.rs:29-56)                                            */
/* ------------------------------------------------------------------------------------------ */

double orc_pair_shift(const orc_pair *p) {
    if (p->eps == 0.0 || p->rc == 0.0) return 0.0;
    double s6 = powi_d(p->sigma / p->rc, 6);
    return 4.0 * p->eps * s6 * (s6 - 1.0);
}

double orc_pair_energy(const orc_pair *p, double r2) {
    if (r2 >= p->rc * p->rc) return 0.0;
    double s6 = powi_d(p->sigma * p->sigma / r2, 3);
    return 4.0 * p->eps * s6 * (s6 - 1.0) - orc_pair_shift(p);
}

double orc_pair_f_over_r(const orc_pair *p, double r2) {
    if (r2 >= p->rc * p->rc) return 0.0;
    double s6 = powi_d(p->sigma * p->sigma / r2, 3);
    return 24.0 * p->eps * s6 * (2.0 * s6 - 1.0) / r2;
}

/* PairConsts<F>::new (crates/md-ref/src/real.rs:115-123) */
typedef struct { float rc2, eps4, sigma2, shift; } consts_f;
typedef struct { double rc2, eps4, sigma2, shift; } consts_d;

static consts_f consts_f32(const orc_pair *p) {
    consts_f c = {(float)(p->rc * p->rc), (float)(4.0 * p->eps), (float)(p->sigma * p->sigma), (float)orc_pair_shift(p)};
    return c;
}

static consts_d consts_f64(const orc_pair *p) {
    consts_d c = {p->rc * p->rc, 4.0 * p->eps, p->sigma * p->sigma, orc_pair_shift(p)};
    return c;
}

/* ------------------------------------------------------------------------------------------ */
/* 3D system helpers (crates/md-core/src/three.rs:103-157)                                    */
/* ------------------------------------------------------------------------------------------ */

static int pinned3(const orc_system3 *s, uint32_t sp) { return (s->pinned >> sp) & 1u; }

static double rc_max3(const orc_system3 *s) {
    double m = 0.0;
    for (uint32_t a = 0; a < s->nspecies; a++)
        for (uint32_t b = 0; b < s->nspecies; b++) m = fmax(m, s->pairs[a * 4 + b].rc);
    return m;
}

static double kick_coeff3(const orc_system3 *s, uint32_t sp) {
    if (pinned3(s, sp)) return 0.0;
    return s->dt * s->dt / s->masses[sp] * ONE_F;
}

static int32_t to_step3(const orc_system3 *s, double v) { return sat_i32(round(v * s->dt * ONE_F)); }
static double velocity3(const orc_system3 *s, int32_t v) { return (double)v * UNIT / s->dt; }

void orc3_cells_for(const orc_system3 *sys, double r, uint32_t out[3]) {
    double b[3], mx = 0.0;
    for (int a = 0; a < 3; a++) {
        b[a] = orc_to_f64(sys->box_len[a]);
        mx = fmax(mx, b[a]);
    }
    double margin = 1e-3 + 1e-6 * mx;
    for (int a = 0; a < 3; a++) out[a] = sat_u32(floor(b[a] / (r + margin)));
}

/* crates/md-ref/src/cells.rs:53-55 */
static uint32_t cell_of(double x, double l, uint32_t n) {
    uint32_t c = sat_u32(x / l * (double)n);
    return c < n - 1 ? c : n - 1;
}

/* ------------------------------------------------------------------------------------------ */
/* 3D init (crates/md-core/src/three.rs:311-404)                                              */
/* ------------------------------------------------------------------------------------------ */

void orc3_cube_for(uint64_t n, double density, int64_t out[3]) {
    double l = cbrt((double)n / density);
    out[0] = out[1] = out[2] = orc_to_fixed(l);
}

void orc3_fcc_cells(uint64_t n, const int64_t box_len[3], uint64_t c[3]) {
    double b[3];
    for (int a = 0; a < 3; a++) b[a] = orc_to_f64(box_len[a]);
    double a = cbrt(b[0] * b[1] * b[2] * 4.0 / (double)n);
    for (int k = 0; k < 3; k++) {
        uint64_t v = sat_u64(round(b[k] / a));
        c[k] = v > 1 ? v : 1;
    }
    while (4 * c[0] * c[1] * c[2] < n) {
        /* the first axis with the widest unit cell */
        int k = 0;
        for (int x = 1; x < 3; x++)
            if (b[x] / (double)c[x] > b[k] / (double)c[k]) k = x;
        c[k] += 1;
    }
}

void orc3_fcc(uint64_t n, const int64_t box_len[3], const uint64_t cells[3], double jitter, uint64_t seed, int64_t *pos) {
    static const double basis[4][3] = {{0.25, 0.25, 0.25}, {0.75, 0.75, 0.25}, {0.75, 0.25, 0.75}, {0.25, 0.75, 0.75}};
    double b[3], a[3];
    for (int x = 0; x < 3; x++) {
        b[x] = orc_to_f64(box_len[x]);
        a[x] = b[x] / (double)cells[x];
    }
    uint64_t id = 0;
    for (uint64_t k = 0; k < cells[2]; k++)
        for (uint64_t j = 0; j < cells[1]; j++)
            for (uint64_t i = 0; i < cells[0]; i++)
                for (int s = 0; s < 4; s++) {
                    if (id == n) return;
                    double cell[3] = {(double)i, (double)j, (double)k};
                    for (int ax = 0; ax < 3; ax++) {
                        double jit = jitter * (orc_uniform(seed, STREAM_JITTER, (uint32_t)(3 * id + ax)) - 0.5);
                        pos[3 * id + ax] = orc_wrap(orc_to_fixed((cell[ax] + basis[s][ax]) * a[ax] + jit), box_len[ax]);
                    }
                    id++;
                }
}

void orc3_zero_velocity_sum(const orc_system3 *sys, uint64_t n, const uint32_t *species, int32_t *vel) {
    int64_t count = 0;
    for (uint64_t i = 0; i < n; i++) count += !pinned3(sys, species[i]);
    if (count == 0) return;
    for (int a = 0; a < 3; a++) {
        int64_t sum = 0;
        for (uint64_t i = 0; i < n; i++)
            if (!pinned3(sys, species[i])) sum += vel[3 * i + a];
        /* Euclidean division */
        int64_t q = sum / count, r = sum % count;
        if (r < 0) {
            r += count;
            q -= 1;
        }
        int64_t k = 0;
        for (uint64_t i = 0; i < n; i++) {
            if (pinned3(sys, species[i])) continue;
            vel[3 * i + a] = (int32_t)((int64_t)vel[3 * i + a] - (q + (k < r ? 1 : 0)));
            k++;
        }
    }
}

void orc3_maxwell(const orc_system3 *sys, uint64_t n, const uint32_t *species, double t, uint64_t seed, int32_t *vel) {
    for (uint64_t i = 0; i < n; i++) {
        uint32_t s = species[i];
        if (pinned3(sys, s)) {
            vel[3 * i] = vel[3 * i + 1] = vel[3 * i + 2] = 0;
            continue;
        }
        double sd = sqrt(t / sys->masses[s]);
        for (int a = 0; a < 3; a++) vel[3 * i + a] = to_step3(sys, sd * orc_normal(seed, STREAM_VELOCITY, (uint32_t)(3 * i + a)));
    }
    orc3_zero_velocity_sum(sys, n, species, vel);
}

void orc3_set_temperature(const orc_system3 *sys, uint64_t n, const uint32_t *species, int32_t *vel, double t) {
    uint64_t moving = 0;
    double ke = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        if (pinned3(sys, species[i])) continue;
        moving++;
        double m = sys->masses[species[i]];
        double s = 0.0;
        for (int a = 0; a < 3; a++) s += powi_d(velocity3(sys, vel[3 * i + a]), 2);
        ke += 0.5 * m * s;
    }
    double dof = 1.5 * fmax((double)moving - 1.0, 1.0);
    double now = ke / dof;
    if (now <= 0.0) return;
    double f = sqrt(t / now);
    for (uint64_t i = 0; i < 3 * n; i++) vel[i] = sat_i32(round((double)vel[i] * f));
    orc3_zero_velocity_sum(sys, n, species, vel);
}

/* three.rs:280-297 */
uint64_t orc3_hash(uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species) {
    uint64_t h = 0xcbf29ce484222325ULL;
#define EAT(v, nb)                                           \
    do {                                                     \
        uint64_t _v = (uint64_t)(v);                         \
        for (int _b = 0; _b < (nb); _b++) {                  \
            h ^= (_v >> (8 * _b)) & 0xff;                    \
            h *= 0x100000001b3ULL;                           \
        }                                                    \
    } while (0)
    for (uint64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            EAT(pos[3 * i + a], 8);
            EAT((uint32_t)vel[3 * i + a], 4);
        }
        EAT(species[i], 4);
    }
    return h;
}

/* ------------------------------------------------------------------------------------------ */
/* IntEngine3<F>::forces (crates/md-ref/src/int_engine3.rs:48-131)                             */
/* ------------------------------------------------------------------------------------------ */

static int64_t separation3(const orc_system3 *s, int64_t a, int64_t b, int axis) {
    return s->walled[axis] ? b - a : orc_min_image(a, b, s->box_len[axis]);
}

typedef struct {
    uint32_t nc[3];
    uint64_t *start; /* ncell + 1 */
    uint64_t *items;
    uint32_t *cell;  /* per particle, 3 words */
} cells3;

static int build_cells3(const orc_system3 *s, uint64_t n, const int64_t *pos, cells3 *c) {
    orc3_cells_for(s, rc_max3(s), c->nc);
    if (c->nc[0] < 3 || c->nc[1] < 3 || c->nc[2] < 3) return -1;
    uint64_t ncell = (uint64_t)c->nc[0] * c->nc[1] * c->nc[2];
    c->start = calloc(ncell + 1, sizeof(uint64_t));
    c->items = malloc((n ? n : 1) * sizeof(uint64_t));
    c->cell = malloc((n ? n : 1) * 3 * sizeof(uint32_t));
    uint64_t *keys = malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t *fill = malloc((ncell + 1) * sizeof(uint64_t));
    if (!c->start || !c->items || !c->cell || !keys || !fill) return -2;
    double b[3];
    for (int a = 0; a < 3; a++) b[a] = orc_to_f64(s->box_len[a]);
    for (uint64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) c->cell[3 * i + a] = cell_of(orc_to_f64(pos[3 * i + a]), b[a], c->nc[a]);
        keys[i] = ((uint64_t)c->cell[3 * i + 2] * c->nc[1] + c->cell[3 * i + 1]) * c->nc[0] + c->cell[3 * i];
        c->start[keys[i] + 1] += 1;
    }
    for (uint64_t k = 0; k < ncell; k++) c->start[k + 1] += c->start[k];
    memcpy(fill, c->start, (ncell + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) c->items[fill[keys[i]]++] = i;
    free(keys);
    free(fill);
    return 0;
}

static void free_cells3(cells3 *c) {
    free(c->start);
    free(c->items);
    free(c->cell);
}

/* The per-pair body and the soft walls in F = float and F = double.  The operation order is
 * PairConsts::eval3 (crates/md-ref/src/real.rs:145-156) and the kick loop
 * (int_engine3.rs:94-101). */
#define DEFINE_FORCES3(NAME, F, CONSTS, MKCONSTS, SEP, LIT)                                                  \
    static int NAME(const orc_system3 *s, uint64_t n, const int64_t *pos, const uint32_t *sp, int32_t *kicks, \
                    double *pe_out, double *wall_force, int brute) {                                          \
        CONSTS consts[16];                                                                                    \
        F kick_c[4];                                                                                          \
        for (int k = 0; k < 16; k++) consts[k] = MKCONSTS(&s->pairs[k]);                                      \
        for (uint32_t k = 0; k < 4; k++) kick_c[k] = k < s->nspecies ? (F)kick_coeff3(s, k) : (F)0;          \
        cells3 c = {{0, 0, 0}, 0, 0, 0};                                                                      \
        if (!brute) {                                                                                         \
            int rc_ = build_cells3(s, n, pos, &c);                                                            \
            if (rc_ == -1) {                                                                                  \
                /* fewer than 3 cells: md-ref asserts; every pair is a candidate instead */                  \
                free_cells3(&c);                                                                              \
                memset(&c, 0, sizeof c);                                                                      \
                brute = 1;                                                                                    \
            } else if (rc_) {                                                                                 \
                free_cells3(&c);                                                                              \
                return rc_;                                                                                   \
            }                                                                                                 \
        }                                                                                                     \
        double pe = 0.0;                                                                                      \
        const F one = LIT(1.0), two = LIT(2.0), six = LIT(6.0);                                              \
        for (uint64_t i = 0; i < n; i++) {                                                                    \
            const int64_t *pi = &pos[3 * i];                                                                  \
            uint32_t si = sp[i];                                                                              \
            F cf = kick_c[si];                                                                                \
            int32_t k[3] = {0, 0, 0};                                                                         \
            F pe_i = LIT(0.0);                                                                                \
            uint64_t nblock = brute ? 1 : 27;                                                                 \
            for (uint64_t blk = 0; blk < nblock; blk++) {                                                     \
                uint64_t q0, q1;                                                                              \
                if (brute) {                                                                                  \
                    q0 = 0;                                                                                   \
                    q1 = n;                                                                                   \
                } else {                                                                                      \
                    uint32_t oz = (uint32_t)(blk / 9), oy = (uint32_t)(blk / 3 % 3), ox = (uint32_t)(blk % 3); \
                    uint32_t z = (c.cell[3 * i + 2] + c.nc[2] + oz - 1) % c.nc[2];                            \
                    uint32_t y = (c.cell[3 * i + 1] + c.nc[1] + oy - 1) % c.nc[1];                            \
                    uint32_t x = (c.cell[3 * i] + c.nc[0] + ox - 1) % c.nc[0];                                \
                    uint64_t cell = ((uint64_t)z * c.nc[1] + y) * c.nc[0] + x;                                \
                    q0 = c.start[cell];                                                                       \
                    q1 = c.start[cell + 1];                                                                   \
                }                                                                                             \
                for (uint64_t q = q0; q < q1; q++) {                                                          \
                    uint64_t j = brute ? q : c.items[q];                                                      \
                    if (j == i) continue;                                                                     \
                    const int64_t *pj = &pos[3 * j];                                                          \
                    F f[3];                                                                                   \
                    int in = 1;                                                                               \
                    for (int a = 0; a < 3; a++) in &= SEP(separation3(s, pi[a], pj[a], a), &f[a]);           \
                    if (!in) continue;                                                                        \
                    CONSTS pc = consts[si * 4 + sp[j]];                                                       \
                    F r2 = f[0] * f[0] + f[1] * f[1] + (orc_mutation == 5 ? (F)0 : f[2] * f[2]);             \
                    if (!(r2 < pc.rc2)) continue;                                                             \
                    F s2 = pc.sigma2 / r2;                                                                    \
                    F s6 = s2 * s2 * s2;                                                                      \
                    F fr = six * pc.eps4 * s6 * (two * s6 - (orc_mutation == 2 ? (F)0 : one)) / r2;           \
                    F u = pc.eps4 * s6 * (s6 - one) - (orc_mutation == 4 ? (F)0 : pc.shift);                  \
                    F aa = orc_mutation == 1 ? fr * cf : -fr * cf;                                            \
                    for (int a = 0; a < 3; a++) k[a] = (int32_t)((uint32_t)k[a] + (uint32_t)sat_i32((double)RINT(aa * f[orc_mutation == 3 ? (a + 1) % 3 : a]))); \
                    pe_i += u;                                                                                \
                }                                                                                             \
            }                                                                                                 \
            double wall_pe = 0.0;                                                                             \
            for (int a = 0; a < 3; a++) {                                                                     \
                if (!s->walled[a]) continue;                                                                  \
                for (int side = 0; side < 2; side++) {                                                        \
                    if (s->wall_kind[a][side] != ORC_WALL_SOFT) continue;                                     \
                    int64_t d = side == 0 ? pi[a] : s->box_len[a] - pi[a];                                    \
                    F df;                                                                                     \
                    if (!SEP(d, &df)) continue;                                                               \
                    if (df < (F)SOFT_RC) {                                                                    \
                        F s2 = one / (df * df);                                                               \
                        F s6 = s2 * s2 * s2;                                                                  \
                        F fw = LIT(24.0) * s6 * (two * s6 - one) / df;                                        \
                        int32_t kw = sat_i32((double)RINT(fw * cf));                                          \
                        k[a] = (int32_t)((uint32_t)k[a] + (side == 0 ? (uint32_t)kw : (uint32_t)0 - (uint32_t)kw)); \
                        if (wall_force) wall_force[6 * i + 2 * a + side] += (double)fw;                       \
                        wall_pe += (double)(LIT(4.0) * s6 * (s6 - one) + one);                                \
                    }                                                                                         \
                }                                                                                             \
            }                                                                                                 \
            memcpy(&kicks[3 * i], k, sizeof k);                                                               \
            pe += 0.5 * (double)pe_i + wall_pe;                                                               \
        }                                                                                                     \
        if (pe_out) *pe_out = pe;                                                                             \
        free_cells3(&c);                                                                                      \
        return 0;                                                                                             \
    }

#define LIT_F(x) ((float)(x))
#define LIT_D(x) ((double)(x))
#define RINT rint_generic
static inline double rint_generic_d(double x) { return rint(x); }
static inline float rint_generic_f(float x) { return rintf(x); }
#define rint_generic(x) _Generic((x), float: rint_generic_f, double: rint_generic_d)(x)

DEFINE_FORCES3(forces3_f32, float, consts_f, consts_f32, orc_sep_f32, LIT_F)
DEFINE_FORCES3(forces3_f64, double, consts_d, consts_f64, sep_f64, LIT_D)

static int forces3(int prec, const orc_system3 *s, uint64_t n, const int64_t *pos, const uint32_t *sp, int32_t *kicks, double *pe, double *wf, int brute) {
    if (s->nspecies < 1 || s->nspecies > 4) return -3;
    if (prec == 32) return forces3_f32(s, n, pos, sp, kicks, pe, wf, brute);
    if (prec == 64) return forces3_f64(s, n, pos, sp, kicks, pe, wf, brute);
    return -4;
}

int orc3_kicks(int prec, const orc_system3 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks, double *pe) {
    return forces3(prec, sys, n, pos, species, kicks, pe, NULL, 0);
}

int orc3_kicks_brute(int prec, const orc_system3 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks) {
    return forces3(prec, sys, n, pos, species, kicks, NULL, NULL, 1);
}

/* three.rs:232-243 (platform libm: Rust std's f64::ln/cos/sin) */
static void thermal_velocity3(const orc_system3 *s, uint32_t id, uint32_t clock, double t, double m, double inward, double *vn, double vt[2]) {
    uint32_t k0 = (uint32_t)s->seed, k1 = (uint32_t)(s->seed >> 32) ^ STREAM_WALL;
    uint32_t r[2], r2[2];
    threefry(id, clock, k0, k1, 20, r);
    threefry(id, clock, k0, k1 ^ 1u, 20, r2);
    double u1 = 1.0 - orc_to_unit_f64(r[0], r[1]);
    double u2 = 1.0 - orc_to_unit_f64(r2[0], 0), u3 = orc_to_unit_f64(r2[1], 0);
    double sd = sqrt(t / m);
    *vn = inward * sd * sqrt(-2.0 * log(u1));
    double rad = sd * sqrt(-2.0 * log(u2));
    double a = TAU * u3;
    vt[0] = rad * cos(a);
    vt[1] = rad * sin(a);
}

/* three.rs:247-265 */
static int andersen3(const orc_system3 *s, uint32_t id, uint32_t clock, uint32_t sp, double out[3]) {
    if (!s->has_bath) return 0;
    if (!((s->bath_species >> sp) & 1u) || pinned3(s, sp)) return 0;
    uint32_t k0 = (uint32_t)s->seed, k1 = (uint32_t)(s->seed >> 32) ^ STREAM_THERMO;
    uint32_t r[2], r2[2], r3[2];
    threefry(id, clock, k0, k1, 20, r);
    if (orc_to_unit_f64(r[0], 0) >= s->bath_rate * s->dt) return 0;
    threefry(id, clock, k0, k1 ^ 1u, 20, r2);
    threefry(id, clock, k0, k1 ^ 2u, 20, r3);
    double sd = sqrt(s->bath_t / s->masses[sp]);
    double rad = sqrt(-2.0 * log(1.0 - orc_to_unit_f64(r[1], 0)));
    double a = TAU * orc_to_unit_f64(r2[0], 0);
    double rad3 = sqrt(-2.0 * log(1.0 - orc_to_unit_f64(r3[0], 0)));
    double a3 = TAU * orc_to_unit_f64(r3[1], 0);
    out[0] = sd * rad * cos(a);
    out[1] = sd * rad * sin(a);
    out[2] = sd * rad3 * cos(a3);
    return 1;
}

/* The same draws for keys (id0 + k, clock) with every key redrawn (bath rate x dt above every
 * uniform): out[6k..6k+6) = the Andersen bath's three velocities and a low heat-bath wall's
 * normal and two tangential ones, as integer displacements per step. */
void orc3_draws(uint64_t seed, double t, double mass, double dt, uint32_t id0, uint64_t count, uint32_t clock, int32_t *out) {
    orc_system3 s;
    memset(&s, 0, sizeof s);
    s.seed = seed;
    s.dt = dt;
    s.nspecies = 1;
    s.masses[0] = mass;
    s.has_bath = 1;
    s.bath_species = 1;
    s.bath_t = t;
    s.bath_rate = 2.0 / dt;
    for (uint64_t k = 0; k < count; k++) {
        uint32_t id = id0 + (uint32_t)k;
        double u[3] = {0.0, 0.0, 0.0}, vn, vt[2];
        andersen3(&s, id, clock, 0, u);
        thermal_velocity3(&s, id, clock, t, mass, 1.0, &vn, vt);
        int32_t *o = out + 6 * k;
        o[0] = to_step3(&s, u[0]);
        o[1] = to_step3(&s, u[1]);
        o[2] = to_step3(&s, u[2]);
        o[3] = to_step3(&s, vn);
        o[4] = to_step3(&s, vt[0]);
        o[5] = to_step3(&s, vt[1]);
    }
}

static int32_t gravity_kick(double g, double dt) { return sat_i32(round(g * dt * dt * ONE_F)); }

/* IntEngine3::step (int_engine3.rs:138-206) */
int orc3_step(int prec, const orc_system3 *s, uint64_t n, int64_t *pos, int32_t *vel, const uint32_t *species, uint64_t steps, uint32_t *clock, orc_ledger3 *ledger) {
    int32_t *kicks = malloc((n ? n : 1) * 3 * sizeof(int32_t));
    double *wf = malloc((n ? n : 1) * 6 * sizeof(double));
    if (!kicks || !wf) return -2;
    orc_ledger3 dummy;
    memset(&dummy, 0, sizeof dummy);
    if (!ledger) ledger = &dummy;
    uint32_t clk = clock ? *clock : 0;
    int32_t g[3];
    for (int a = 0; a < 3; a++) g[a] = gravity_kick(s->gravity[a], s->dt);
    int rc = 0;
    for (uint64_t st = 0; st < steps; st++) {
        memset(wf, 0, n * 6 * sizeof(double));
        rc = forces3(prec, s, n, pos, species, kicks, NULL, wf, 0);
        if (rc) break;
        for (uint64_t i = 0; i < n; i++)
            for (int w = 0; w < 6; w++) ledger->impulse[w] += wf[6 * i + w] * s->dt;
        for (uint64_t i = 0; i < n; i++) {
            uint32_t sp = species[i];
            int pin = pinned3(s, sp);
            int32_t v[3];
            for (int a = 0; a < 3; a++) v[a] = (int32_t)((uint32_t)vel[3 * i + a] + (uint32_t)kicks[3 * i + a] + (uint32_t)(pin ? 0 : g[a]));
            double m = s->masses[sp];
            double u[3];
            if (andersen3(s, (uint32_t)i, clk, sp, u)) {
                double before = 0.0, after = 0.0;
                for (int a = 0; a < 3; a++) before += powi_d(velocity3(s, v[a]), 2);
                for (int a = 0; a < 3; a++) v[a] = to_step3(s, u[a]);
                for (int a = 0; a < 3; a++) after += powi_d(velocity3(s, v[a]), 2);
                ledger->energy[6] += 0.5 * m * (after - before);
                ledger->hits[6] += 1;
            }
            int64_t x[3];
            for (int a = 0; a < 3; a++) x[a] = pos[3 * i + a] + (int64_t)v[a];
            for (int a = 0; a < 3; a++) {
                int64_t l = s->box_len[a];
                if (!s->walled[a]) {
                    x[a] = orc_wrap(x[a], l);
                    continue;
                }
                int side = -1;
                if (x[a] < 0) {
                    x[a] = -1 - x[a];
                    side = 0;
                } else if (x[a] >= l) {
                    x[a] = 2 * l - 1 - x[a];
                    side = 1;
                }
                if (side < 0) continue;
                /* collide (int_engine3.rs:186-206) */
                int w = 2 * a + side;
                double before[3], after[3];
                for (int b = 0; b < 3; b++) before[b] = velocity3(s, v[b]);
                if (s->wall_kind[a][side] == ORC_WALL_THERMAL) {
                    double vn, vt[2];
                    thermal_velocity3(s, (uint32_t)i, clk, s->wall_t[a][side], m, side == 0 ? 1.0 : -1.0, &vn, vt);
                    v[a] = to_step3(s, vn);
                    v[(a + 1) % 3] = to_step3(s, vt[0]);
                    v[(a + 2) % 3] = to_step3(s, vt[1]);
                } else {
                    v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
                }
                for (int b = 0; b < 3; b++) after[b] = velocity3(s, v[b]);
                double kea = 0.5 * m * (after[0] * after[0] + after[1] * after[1] + after[2] * after[2]);
                double keb = 0.5 * m * (before[0] * before[0] + before[1] * before[1] + before[2] * before[2]);
                ledger->energy[w] += kea - keb;
                ledger->impulse[w] += fabs(m * (before[a] - after[a]));
                ledger->hits[w] += 1;
            }
            memcpy(&vel[3 * i], v, sizeof v);
            memcpy(&pos[3 * i], x, sizeof x);
        }
        clk += 1;
    }
    if (clock) *clock = clk;
    free(kicks);
    free(wf);
    return rc;
}

/* IntEngine3::reverse (int_engine3.rs:209-219) */
int orc3_reverse(int prec, const orc_system3 *s, uint64_t n, const int64_t *pos, int32_t *vel, const uint32_t *species) {
    int32_t *kicks = malloc((n ? n : 1) * 3 * sizeof(int32_t));
    if (!kicks) return -2;
    int rc = forces3(prec, s, n, pos, species, kicks, NULL, NULL, 0);
    if (!rc) {
        for (uint64_t i = 0; i < n; i++) {
            int pin = pinned3(s, species[i]);
            for (int a = 0; a < 3; a++) {
                uint32_t g = pin ? 0 : (uint32_t)gravity_kick(s->gravity[a], s->dt);
                vel[3 * i + a] = (int32_t)((uint32_t)0 - ((uint32_t)vel[3 * i + a] + (uint32_t)kicks[3 * i + a] + g));
            }
        }
    }
    free(kicks);
    return rc;
}

/* IntEngine3::energy (int_engine3.rs:222-239) */
int orc3_energy(int prec, const orc_system3 *s, uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species, double *kinetic, double *potential) {
    int32_t *kicks = malloc((n ? n : 1) * 3 * sizeof(int32_t));
    if (!kicks) return -2;
    double pe = 0.0;
    int rc = forces3(prec, s, n, pos, species, kicks, &pe, NULL, 0);
    double ke = 0.0;
    if (!rc) {
        for (uint64_t i = 0; i < n; i++) {
            if (pinned3(s, species[i])) continue;
            double m = s->masses[species[i]];
            for (int a = 0; a < 3; a++) {
                int32_t next = (int32_t)((uint32_t)vel[3 * i + a] + (uint32_t)kicks[3 * i + a] + (uint32_t)gravity_kick(s->gravity[a], s->dt));
                double v = 0.5 * (velocity3(s, vel[3 * i + a]) + velocity3(s, next));
                ke += 0.5 * m * v * v;
                pe -= m * s->gravity[a] * orc_to_f64(pos[3 * i + a]);
            }
        }
    }
    *kinetic = ke;
    *potential = pe;
    free(kicks);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* 2D (crates/md-core/src/init.rs, system.rs; crates/md-ref/src/int_engine.rs)                 */
/* ------------------------------------------------------------------------------------------ */

static int pinned2(const orc_system2 *s, uint32_t sp) { return (s->pinned >> sp) & 1u; }

static double kick_coeff2(const orc_system2 *s, uint32_t sp) {
    if (pinned2(s, sp)) return 0.0;
    return s->dt * s->dt / s->masses[sp] * ONE_F;
}

static int32_t to_step2(const orc_system2 *s, double v) { return sat_i32(round(v * s->dt * ONE_F)); }
static double velocity2(const orc_system2 *s, int32_t v) { return (double)v * UNIT / s->dt; }

static double rc_max2(const orc_system2 *s) {
    double m = 0.0;
    for (uint32_t a = 0; a < s->nspecies; a++)
        for (uint32_t b = 0; b < s->nspecies; b++) m = fmax(m, s->pairs[a * 4 + b].rc);
    return m;
}

/* system.rs cells_for (2D) */
void orc2_cells_for(const orc_system2 *sys, double r, uint32_t out[2]) {
    double b0 = orc_to_f64(sys->box_len[0]), b1 = orc_to_f64(sys->box_len[1]);
    double margin = 1e-3 + 1e-6 * fmax(b0, b1);
    out[0] = sat_u32(floor(b0 / (r + margin)));
    out[1] = sat_u32(floor(b1 / (r + margin)));
}

/* init.rs:15-19 */
void orc2_box_for(uint64_t n, double density, double aspect, int64_t out[2]) {
    double area = (double)n / density;
    double ly = sqrt(area / aspect);
    out[0] = orc_to_fixed(ly * aspect);
    out[1] = orc_to_fixed(ly);
}

/* init.rs:22-38 */
void orc2_lattice(uint64_t n, const int64_t box_len[2], double jitter, uint64_t seed, int64_t *pos) {
    double b0 = orc_to_f64(box_len[0]), b1 = orc_to_f64(box_len[1]);
    uint64_t nx = sat_u64(ceil(sqrt((double)n * b0 / b1)));
    if (nx < 1) nx = 1;
    uint64_t ny = (n + nx - 1) / nx;
    double a0 = b0 / (double)nx, a1 = b1 / (double)ny;
    for (uint64_t k = 0; k < n; k++) {
        uint64_t i = k % nx, j = k / nx;
        double jx = jitter * (orc_uniform(seed, STREAM_JITTER, (uint32_t)(2 * k)) - 0.5);
        double jy = jitter * (orc_uniform(seed, STREAM_JITTER, (uint32_t)(2 * k + 1)) - 0.5);
        pos[2 * k] = orc_wrap(orc_to_fixed(((double)i + 0.5) * a0 + jx), box_len[0]);
        pos[2 * k + 1] = orc_wrap(orc_to_fixed(((double)j + 0.5) * a1 + jy), box_len[1]);
    }
}

/* init.rs:76-89 (2D: over every particle) */
static void zero_velocity_sum2(uint64_t n, int32_t *vel) {
    if (n == 0) return;
    int64_t nn = (int64_t)n;
    for (int a = 0; a < 2; a++) {
        int64_t sum = 0;
        for (uint64_t i = 0; i < n; i++) sum += vel[2 * i + a];
        int64_t q = sum / nn, r = sum % nn;
        if (r < 0) {
            r += nn;
            q -= 1;
        }
        for (uint64_t i = 0; i < n; i++) vel[2 * i + a] = (int32_t)((int64_t)vel[2 * i + a] - (q + ((int64_t)i < r ? 1 : 0)));
    }
}

/* init.rs:42-57 */
void orc2_maxwell(const orc_system2 *sys, uint64_t n, const uint32_t *species, double t, uint64_t seed, int32_t *vel) {
    for (uint64_t i = 0; i < n; i++) {
        double sd = sqrt(t / sys->masses[species[i]]);
        vel[2 * i] = to_step2(sys, sd * orc_normal(seed, STREAM_VELOCITY, (uint32_t)(2 * i)));
        vel[2 * i + 1] = to_step2(sys, sd * orc_normal(seed, STREAM_VELOCITY, (uint32_t)(2 * i + 1)));
    }
    zero_velocity_sum2(n, vel);
}

/* init.rs:107-121 */
void orc2_fluid(const orc_system2 *sys, uint64_t n, double t, uint64_t seed, int64_t *pos, int32_t *vel, uint32_t *species) {
    const double margin = 1.0;
    int64_t inner[2];
    for (int a = 0; a < 2; a++) inner[a] = sys->walled[a] ? sys->box_len[a] - orc_to_fixed(2.0 * margin) : sys->box_len[a];
    orc2_lattice(n, inner, 0.1, seed, pos);
    for (uint64_t i = 0; i < n; i++)
        for (int a = 0; a < 2; a++)
            if (sys->walled[a]) pos[2 * i + a] += orc_to_fixed(margin);
    for (uint64_t i = 0; i < n; i++) species[i] = 0;
    orc2_maxwell(sys, n, species, t, seed, vel);
}

/* system.rs:322-339 */
uint64_t orc2_hash(uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species) {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (uint64_t i = 0; i < n; i++) {
        EAT(pos[2 * i], 8);
        EAT(pos[2 * i + 1], 8);
        EAT((uint32_t)vel[2 * i], 4);
        EAT((uint32_t)vel[2 * i + 1], 4);
        EAT(species[i], 4);
    }
    return h;
}

/* ---- IntTables (crates/md-core/src/table.rs:45-105) ---- */
#define SEP_BITS 20
#define R2_BITS 24
#define T_BITS 2
#define KICK_SHIFT (T_BITS + SEP_BITS)
#define STEP_BITS 12

uint64_t orc2_int_table_len(const orc_system2 *sys) {
    double rc2_max = rc_max2(sys) * rc_max2(sys);
    return (sat_u64(rc2_max * (double)(1ULL << R2_BITS)) >> STEP_BITS) + 2;
}

void orc2_int_tables(const orc_system2 *sys, int32_t *data, uint32_t *rc2) {
    uint32_t ns = sys->nspecies;
    uint64_t len = orc2_int_table_len(sys);
    for (uint32_t a = 0; a < ns; a++) {
        double c = kick_coeff2(sys, a);
        for (uint32_t b = 0; b < ns; b++) {
            const orc_pair *p = &sys->pairs[a * 4 + b];
            rc2[a * ns + b] = sat_u32(p->rc * p->rc * (double)(1ULL << R2_BITS));
            orc_pair cont = *p;
            cont.rc = INFINITY;
            for (uint64_t k = 0; k < len; k++) {
                double r2 = (double)(k * (1ULL << STEP_BITS)) / (double)(1ULL << R2_BITS);
                r2 = fmax(r2, 0.36);
                double t = 0.0;
                if (r2 < p->rc * p->rc + 2.0 * (double)(1ULL << STEP_BITS) / (double)(1ULL << R2_BITS))
                    t = orc_pair_f_over_r(&cont, r2) * c * (double)(1ULL << T_BITS);
                double rt = round(t);
                if (rt < -2147483648.0) rt = -2147483648.0;
                if (rt > 2147483647.0) rt = 2147483647.0;
                data[((uint64_t)a * ns + b) * len + k] = sat_i32(rt);
            }
        }
    }
}

int orc2_int_kick(const orc_system2 *sys, const int32_t *data, const uint32_t *rc2, uint64_t len, uint32_t sa, uint32_t sb, int64_t dx, int64_t dy, int32_t out[2]) {
    const int64_t limit = 8LL << 32;
    int64_t d[2] = {dx, dy}, q[2];
    for (int a = 0; a < 2; a++)
        if (!(d[a] >= -limit && d[a] < limit)) return 0;
    for (int a = 0; a < 2; a++) {
        uint64_t m = d[a] < 0 ? (uint64_t)0 - (uint64_t)d[a] : (uint64_t)d[a];
        int64_t mm = (int64_t)((m + (1ULL << (31 - SEP_BITS))) >> (32 - SEP_BITS));
        q[a] = d[a] < 0 ? -mm : mm;
    }
    uint32_t r2 = (uint32_t)(((uint64_t)(q[0] * q[0] + q[1] * q[1]) + (1ULL << (2 * SEP_BITS - R2_BITS - 1))) >> (2 * SEP_BITS - R2_BITS));
    uint64_t pair = (uint64_t)sa * sys->nspecies + sb;
    if (r2 >= rc2[pair]) return 0;
    uint64_t idx = r2 >> STEP_BITS;
    int64_t frac = (int64_t)(r2 & ((1u << STEP_BITS) - 1));
    uint64_t base = pair * len;
    int64_t t0 = data[base + idx], t1 = data[base + idx + 1];
    int64_t t = t0 + (((t1 - t0) * frac) >> STEP_BITS);
    for (int a = 0; a < 2; a++) {
        unsigned __int128 prod = (unsigned __int128)(t < 0 ? (uint64_t)0 - (uint64_t)t : (uint64_t)t) *
                                 (unsigned __int128)(q[a] < 0 ? (uint64_t)0 - (uint64_t)q[a] : (uint64_t)q[a]);
        int64_t m = (int64_t)(uint64_t)(prod >> KICK_SHIFT);
        int neg = (t < 0) != (q[a] < 0);
        out[a] = (int32_t)(uint32_t)(uint64_t)(neg ? m : -m);
    }
    return 1;
}

/* ---- 2D forces and step: IntEngine (crates/md-ref/src/int_engine.rs:130-400) with the full
 * md_core::System boundaries (system.rs, walls.rs) over the stable cell list
 * (crates/md-ref/src/cells.rs).  prec 0 selects the integer tables. ---- */
#define WALLS2 7
#define MEMBRANE_LOW 4
#define MEMBRANE_HIGH 5
#define BATH2 6
#define MAX_BONDS 4
#define PISTON_FORCE_SCALE 65536.0

static int64_t separation2(const orc_system2 *s, int64_t a, int64_t b, int axis) {
    return s->walled[axis] ? b - a : orc_min_image(a, b, s->box_len[axis]);
}

/* system.rs:253-257, 229-231 */
static int64_t compartment_of2(const orc_system2 *s, int64_t p, int a) { return s->compartments[a] ? p >> s->compartments[a] : 0; }
static int interact_across2(const orc_system2 *s, const int64_t *a, const int64_t *b) {
    return s->diathermal || (compartment_of2(s, a[0], 0) == compartment_of2(s, b[0], 0) && compartment_of2(s, a[1], 1) == compartment_of2(s, b[1], 1));
}

/* cells for the species present (int_engine.rs:108-116) */
static void present_cells2(const orc_system2 *s, uint64_t n, const uint32_t *species, uint32_t nc[2]) {
    uint32_t present = 0;
    for (uint64_t i = 0; i < n; i++) present |= 1u << species[i];
    orc2_cells_for(s, rc_max2(s), nc);
    double rc = 0.0;
    for (uint32_t a = 0; a < s->nspecies; a++)
        for (uint32_t b = 0; b < s->nspecies; b++)
            if ((present >> a & 1u) && (present >> b & 1u)) rc = fmax(rc, s->pairs[a * 4 + b].rc);
    if (rc <= 0.0) return;
    uint32_t c[2];
    orc2_cells_for(s, rc, c);
    nc[0] = c[0] > nc[0] ? c[0] : nc[0];
    nc[1] = c[1] > nc[1] ? c[1] : nc[1];
}

/* fixed::sep_long_f32 (fixed.rs:65-72) */
static float sep_long_f32(int64_t d) {
    float f;
    if (orc_sep_f32(d, &f)) return f;
    uint64_t m = d < 0 ? (uint64_t)0 - (uint64_t)d : (uint64_t)d;
    float mag = (float)(uint32_t)(m >> 32) + (float)(uint32_t)m * (1.0f / 4294967296.0f);
    return d < 0 ? -mag : mag;
}

#define DEFINE_FORCES2(NAME, F, CONSTS, MKCONSTS, SEP, LIT)                                                  \
    static void NAME(const orc_system2 *s, uint64_t i, const int64_t *pos, const uint32_t *sp, uint64_t j,    \
                     const CONSTS *consts, F cf, int32_t k[2], F *pe_i) {                                      \
        const int64_t *pi = &pos[2 * i], *pj = &pos[2 * j];                                                    \
        F fx, fy;                                                                                              \
        if (!SEP(separation2(s, pi[0], pj[0], 0), &fx)) return;                                                \
        if (!SEP(separation2(s, pi[1], pj[1], 1), &fy)) return;                                                \
        CONSTS pc = consts[sp[i] * 4 + sp[j]];                                                                 \
        F r2 = fx * fx + fy * fy;                                                                              \
        if (!(r2 < pc.rc2)) return;                                                                            \
        const F one = LIT(1.0);                                                                                \
        F s2 = pc.sigma2 / r2;                                                                                 \
        F s6 = s2 * s2 * s2;                                                                                   \
        F fr = LIT(6.0) * pc.eps4 * s6 * (LIT(2.0) * s6 - one) / r2;                                           \
        F u = pc.eps4 * s6 * (s6 - one) - pc.shift;                                                           \
        F aa = -fr * cf;                                                                                       \
        k[0] = (int32_t)((uint32_t)k[0] + (uint32_t)sat_i32((double)RINT(aa * fx)));                          \
        k[1] = (int32_t)((uint32_t)k[1] + (uint32_t)sat_i32((double)RINT(aa * fy)));                          \
        *pe_i += u;                                                                                            \
    }

DEFINE_FORCES2(pair2_f32, float, consts_f, consts_f32, orc_sep_f32, LIT_F)
DEFINE_FORCES2(pair2_f64, double, consts_d, consts_f64, sep_f64, LIT_D)

/* The per-particle terms after the pairs (int_engine.rs:181-248): bonds (BondConsts::eval,
 * real.rs:167-196, on sep_long separations), soft walls, and the piston as a soft wall at
 * height piston_w pushing down.  wf: the particle's force on each wall (outward). */
#define DEFINE_EXTRAS2(NAME, F, SEP, SEPLONG, LIT)                                                           \
    static void NAME(const orc_system2 *s, const orc_state2 *st, uint64_t i, const int64_t *pos,             \
                     const uint32_t *partners, F cf, int32_t k[2], double *pe, double *wf, int64_t *piston_fixed) { \
        const int64_t *pi = &pos[2 * i];                                                                      \
        const F one = LIT(1.0), half = LIT(0.5);                                                             \
        F bond_pe = LIT(0.0);                                                                                 \
        for (int b = 0; b < MAX_BONDS; b++) {                                                                 \
            uint32_t j = partners[8 * i + 2 * b];                                                             \
            if (j == UINT32_MAX) break;                                                                       \
            uint32_t t = partners[8 * i + 2 * b + 1];                                                         \
            const int64_t *pj = &pos[2 * (uint64_t)j];                                                        \
            F fx = SEPLONG(separation2(s, pi[0], pj[0], 0)), fy = SEPLONG(separation2(s, pi[1], pj[1], 1));  \
            F r2 = fx * fx + fy * fy;                                                                         \
            F bk = (F)s->bond_k[t], br0 = (F)s->bond_r0[t], fr, u;                                            \
            if (s->bond_fene[t]) {                                                                            \
                F x = r2 / (br0 * br0);                                                                       \
                F cap = (F)0.999999;                                                                          \
                if (x > cap) x = cap;                                                                         \
                u = -half * bk * br0 * br0 * (F)log((double)(one - x));                                       \
                fr = -bk / (one - x);                                                                         \
            } else if (br0 == LIT(0.0)) {                                                                     \
                fr = -bk;                                                                                     \
                u = half * bk * r2;                                                                           \
            } else {                                                                                          \
                F r = (F)sqrt((double)r2);                                                                    \
                F dr = r - br0;                                                                               \
                fr = -bk * dr / r;                                                                            \
                u = half * bk * dr * dr;                                                                      \
            }                                                                                                 \
            F aa = -fr * cf;                                                                                  \
            k[0] = (int32_t)((uint32_t)k[0] + (uint32_t)sat_i32((double)RINT(aa * fx)));                     \
            k[1] = (int32_t)((uint32_t)k[1] + (uint32_t)sat_i32((double)RINT(aa * fy)));                     \
            bond_pe += u;                                                                                     \
        }                                                                                                     \
        *pe += 0.5 * (double)bond_pe;                                                                         \
        for (int a = 0; a < 2; a++) {                                                                         \
            if (!s->walled[a]) continue;                                                                      \
            for (int side = 0; side < 2; side++) {                                                            \
                if (s->wall_kind[a][side] != ORC_WALL_SOFT) continue;                                         \
                int64_t d = side == 0 ? pi[a] : s->box_len[a] - pi[a];                                        \
                F df;                                                                                         \
                if (!SEP(d, &df)) continue;                                                                   \
                if (df < (F)SOFT_RC) {                                                                        \
                    F s2 = one / (df * df);                                                                   \
                    F s6 = s2 * s2 * s2;                                                                      \
                    F f = LIT(24.0) * s6 * (LIT(2.0) * s6 - one) / df;                                        \
                    int32_t kw = sat_i32((double)RINT(f * cf));                                               \
                    k[a] = (int32_t)((uint32_t)k[a] + (side == 0 ? (uint32_t)kw : (uint32_t)0 - (uint32_t)kw)); \
                    if (wf) wf[WALLS2 * i + 2 * a + side] += (double)f;                                       \
                    *pe += (double)(LIT(4.0) * s6 * (s6 - one) + one);                                        \
                }                                                                                             \
            }                                                                                                 \
        }                                                                                                     \
        if (s->has_piston) {                                                                                  \
            F df;                                                                                             \
            if (SEP(st->piston_w - pi[1], &df) && df < (F)SOFT_RC) {                                          \
                F s2 = one / (df * df);                                                                       \
                F s6 = s2 * s2 * s2;                                                                          \
                F f = LIT(24.0) * s6 * (LIT(2.0) * s6 - one) / df;                                            \
                k[1] = (int32_t)((uint32_t)k[1] - (uint32_t)sat_i32((double)RINT(f * cf)));                  \
                *piston_fixed += (int64_t)round((double)f * PISTON_FORCE_SCALE);                              \
                *pe += (double)(LIT(4.0) * s6 * (s6 - one) + one);                                            \
            }                                                                                                 \
        }                                                                                                     \
    }

static double sep_long_f64(int64_t d) { return orc_to_f64(d); }
DEFINE_EXTRAS2(extras2_f32, float, orc_sep_f32, sep_long_f32, LIT_F)
DEFINE_EXTRAS2(extras2_f64, double, sep_f64, sep_long_f64, LIT_D)

/* System::partner_table (system.rs:312-324): per particle id, up to 4 (partner id, type). */
static uint32_t *partner_table2(const orc_system2 *s, uint64_t n) {
    uint32_t *t = malloc((n ? n : 1) * 8 * sizeof(uint32_t));
    uint8_t *used = calloc(n ? n : 1, 1);
    if (!t || !used) {
        free(t);
        free(used);
        return NULL;
    }
    for (uint64_t k = 0; k < n * 8; k++) t[k] = UINT32_MAX;
    for (uint64_t b = 0; b < s->nbonds; b++) {
        const uint32_t *bd = &s->bonds[3 * b];
        for (int e = 0; e < 2; e++) {
            uint32_t me = bd[e], other = bd[1 - e];
            if (me >= n || used[me] >= MAX_BONDS) continue;
            t[8 * (uint64_t)me + 2 * used[me]] = other;
            t[8 * (uint64_t)me + 2 * used[me] + 1] = bd[2];
            used[me]++;
        }
    }
    free(used);
    return t;
}

void orc2_state_init(const orc_system2 *s, orc_state2 *st) {
    memset(st, 0, sizeof *st);
    if (s->has_piston) {
        st->piston_w = orc_to_fixed(s->piston_start);
        st->piston_v = s->piston_free ? 0 : to_step2(s, s->piston_speed);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* md-ref's fast 2D pair path (crates/md-ref/src/fast.rs), the path IntEngine takes by default   */
/* (fast: true, int_engine.rs:89; chosen at int_engine.rs:136-141): a counting sort into cells,   */
/* a half shell of neighbour cells so each pair is visited once, and per cell a branch-free row   */
/* per particle over the cell's neighbourhood.  Kicks are bit-identical to the reference loop    */
/* (each end's kick is rounded with its own coefficient from the same force factor, and integer  */
/* sums do not depend on order); the potential energy is summed in another order.                */
/* ------------------------------------------------------------------------------------------ */

int orc2_fast_path = 0;
void orc2_set_fast_path(int on) { orc2_fast_path = on; }

/* fast::eligible (fast.rs:76-80) */
static int fast_eligible2(const orc_system2 *s, const uint32_t nc[2]) {
    double side = fmax(orc_to_f64(s->box_len[0]) / nc[0], orc_to_f64(s->box_len[1]) / nc[1]);
    return (!(s->compartments[0] || s->compartments[1]) || s->diathermal) && 2.0 * side < 7.9;
}

/* `x as i32` after round_ties_even, without branches (the row loops below stay vectorisable):
 * NaN -> 0, saturating at the i32 range.  The same values as sat_i32. */
static inline int32_t sat_i32_rf(float x) {
    const float c = fmaxf(fminf(x, 2147483520.0f), -2147483648.0f);  /* the largest f32 below 2^31 */
    int32_t r = (int32_t)c;
    r = x >= 2147483648.0f ? INT32_MAX : r;
    return x != x ? 0 : r;
}
static inline int32_t sat_i32_rd(double x) {
    const double c = fmax(fmin(x, 2147483647.0), -2147483648.0);
    int32_t r = (int32_t)c;
    return x != x ? 0 : r;
}
#define SAT_I32_R(x) _Generic((x), float: sat_i32_rf, double: sat_i32_rd)(x)

/* Real::sep_near (real.rs:59-64 for f32; to_f64 for f64): the separation conversion for pairs
 * known to lie within 8 sigma. */
static float sep_near_f32(int64_t d) {
    float f;
    orc_sep_f32(d, &f);
    return f;
}
static double sep_near_f64(int64_t d) { return orc_to_f64(d); }

/* fast::pair_kicks (fast.rs:83-237) with row_uniform / row_mixture (:239-294).  kicks (by id)
 * receive the pair kicks; returns the pairs' potential energy in F. */
#define DEFINE_FAST2(NAME, F, CONSTS, SEPNEAR, LIT)                                                          \
    static F NAME(const orc_system2 *s, uint64_t n, const int64_t *pos, const uint32_t *sp, const CONSTS *consts, \
                  const F *kc, const uint32_t nc[2], int32_t *kicks, int *err) {                               \
        const uint32_t cnx = nc[0], cny = nc[1], ns = s->nspecies;                                            \
        const uint64_t ncell = (uint64_t)cnx * cny;                                                           \
        const double b0 = orc_to_f64(s->box_len[0]), b1 = orc_to_f64(s->box_len[1]);                          \
        uint32_t *start = calloc(ncell + 1, sizeof(uint32_t)), *fillp = malloc((ncell + 1) * sizeof(uint32_t)); \
        uint32_t *key = malloc((n ? n : 1) * sizeof(uint32_t)), *ss = malloc((n ? n : 1) * sizeof(uint32_t));  \
        uint32_t *sid = malloc((n ? n : 1) * sizeof(uint32_t));                                               \
        int64_t *sx = malloc((n ? n : 1) * sizeof(int64_t)), *sy = malloc((n ? n : 1) * sizeof(int64_t));     \
        int32_t *kx = calloc(n ? n : 1, sizeof(int32_t)), *ky = calloc(n ? n : 1, sizeof(int32_t));           \
        const uint64_t cap = n ? n : 1; /* a neighbourhood holds at most every particle */                   \
        int64_t *nx = malloc(cap * sizeof(int64_t)), *ny = malloc(cap * sizeof(int64_t));                      \
        uint32_t *ni = malloc(cap * sizeof(uint32_t));                                                        \
        int32_t *nkx = malloc(cap * sizeof(int32_t)), *nky = malloc(cap * sizeof(int32_t));                  \
        F *rx = malloc(cap * sizeof(F)), *ry = malloc(cap * sizeof(F));                                       \
        F pe = LIT(0.0);                                                                                      \
        *err = 0;                                                                                             \
        if (!start || !fillp || !key || !ss || !sid || !sx || !sy || !kx || !ky || !nx || !ny || !ni || !nkx || !nky || !rx || !ry) { \
            *err = -2;                                                                                        \
            goto done;                                                                                        \
        }                                                                                                     \
        /* counting sort into cells (cells.rs:51-55 cell_of, y-major keys) */                                 \
        for (uint64_t i = 0; i < n; i++) {                                                                    \
            uint32_t cx = cell_of(orc_to_f64(pos[2 * i]), b0, cnx), cy = cell_of(orc_to_f64(pos[2 * i + 1]), b1, cny); \
            key[i] = cy * cnx + cx;                                                                           \
            start[key[i] + 1] += 1;                                                                           \
        }                                                                                                     \
        for (uint64_t c = 0; c < ncell; c++) start[c + 1] += start[c];                                        \
        memcpy(fillp, start, (ncell + 1) * sizeof(uint32_t));                                                 \
        for (uint64_t i = 0; i < n; i++) {                                                                    \
            uint32_t q = fillp[key[i]]++;                                                                     \
            sx[q] = pos[2 * i];                                                                               \
            sy[q] = pos[2 * i + 1];                                                                           \
            ss[q] = sp[i];                                                                                    \
            sid[q] = (uint32_t)i;                                                                             \
        }                                                                                                     \
        /* the half shell: the cell itself (later particles) and four of its eight neighbours */              \
        static const int SHELL[4][2] = {{1, 0}, {-1, 1}, {0, 1}, {1, 1}};                                      \
        const F zero = LIT(0.0), one = LIT(1.0), two = LIT(2.0), six = LIT(6.0);                              \
        for (uint32_t cy = 0; cy < cny; cy++)                                                                 \
            for (uint32_t cx = 0; cx < cnx; cx++) {                                                           \
                const uint32_t c = cy * cnx + cx, a0 = start[c], a1 = start[c + 1];                           \
                if (a0 == a1) continue;                                                                       \
                uint64_t m = 0;                                                                               \
                /* gather the neighbourhood: own cell first, then the shell, shifted across periodic edges */ \
                for (int r = -1; r < 4; r++) {                                                                \
                    uint32_t d = c;                                                                           \
                    int64_t shx = 0, shy = 0;                                                                 \
                    if (r >= 0) {                                                                             \
                        int tx = (int)cx + SHELL[r][0], ty = (int)cy + SHELL[r][1];                           \
                        if (tx < 0) {                                                                         \
                            if (s->walled[0]) continue;                                                       \
                            tx += (int)cnx;                                                                   \
                            shx = -s->box_len[0];                                                             \
                        } else if (tx >= (int)cnx) {                                                          \
                            if (s->walled[0]) continue;                                                       \
                            tx -= (int)cnx;                                                                   \
                            shx = s->box_len[0];                                                              \
                        }                                                                                     \
                        if (ty >= (int)cny) {                                                                 \
                            if (s->walled[1]) continue;                                                       \
                            ty -= (int)cny;                                                                   \
                            shy = s->box_len[1];                                                              \
                        }                                                                                     \
                        d = (uint32_t)ty * cnx + (uint32_t)tx;                                                \
                    }                                                                                         \
                    for (uint32_t q = start[d]; q < start[d + 1]; q++) {                                      \
                        nx[m] = sx[q] + shx;                                                                  \
                        ny[m] = sy[q] + shy;                                                                  \
                        ni[m] = q;                                                                            \
                        m++;                                                                                  \
                    }                                                                                         \
                }                                                                                             \
                memset(nkx, 0, m * sizeof(int32_t));                                                          \
                memset(nky, 0, m * sizeof(int32_t));                                                          \
                for (uint32_t p = 0; p < a1 - a0; p++) {                                                      \
                    const uint32_t sq = a0 + p;                                                               \
                    const int64_t xs = sx[sq], ys = sy[sq];                                                   \
                    for (uint64_t k = 0; k < m; k++) {                                                        \
                        rx[k] = SEPNEAR(nx[k] - xs);                                                          \
                        ry[k] = SEPNEAR(ny[k] - ys);                                                          \
                    }                                                                                         \
                    const uint32_t si = ss[sq];                                                               \
                    const F ci = kc[si];                                                                      \
                    int32_t kxi = 0, kyi = 0;                                                                 \
                    for (uint64_t k = 0; k < m; k++) {                                                        \
                        const uint32_t sj = ns == 1 ? 0u : ss[ni[k]];                                         \
                        const CONSTS pc = consts[si * 4 + sj];                                                \
                        const F x = rx[k], y = ry[k];                                                         \
                        const F r2 = x * x + y * y;                                                           \
                        const F s2 = pc.sigma2 / r2;                                                          \
                        const F s6 = s2 * s2 * s2;                                                            \
                        const F f = six * pc.eps4 * s6 * (two * s6 - one) / r2;                               \
                        const F e = pc.eps4 * s6 * (s6 - one) - pc.shift;                                     \
                        const int on = r2 < pc.rc2 && k >= (uint64_t)p + 1;                                   \
                        const F fo = on ? f : zero;                                                           \
                        const F ai = -fo * ci, aj = -fo * kc[sj];                                             \
                        kxi = (int32_t)((uint32_t)kxi + (uint32_t)SAT_I32_R(RINT(ai * x)));                   \
                        kyi = (int32_t)((uint32_t)kyi + (uint32_t)SAT_I32_R(RINT(ai * y)));                   \
                        nkx[k] = (int32_t)((uint32_t)nkx[k] + (uint32_t)SAT_I32_R(RINT(aj * -x)));            \
                        nky[k] = (int32_t)((uint32_t)nky[k] + (uint32_t)SAT_I32_R(RINT(aj * -y)));            \
                        pe += on ? e : zero;                                                                  \
                    }                                                                                         \
                    kx[sq] = (int32_t)((uint32_t)kx[sq] + (uint32_t)kxi);                                    \
                    ky[sq] = (int32_t)((uint32_t)ky[sq] + (uint32_t)kyi);                                    \
                }                                                                                             \
                /* scatter the neighbourhood's kicks back, once per cell */                                   \
                for (uint64_t k = 0; k < m; k++) {                                                            \
                    kx[ni[k]] = (int32_t)((uint32_t)kx[ni[k]] + (uint32_t)nkx[k]);                           \
                    ky[ni[k]] = (int32_t)((uint32_t)ky[ni[k]] + (uint32_t)nky[k]);                           \
                }                                                                                             \
            }                                                                                                 \
        for (uint64_t q = 0; q < n; q++) {                                                                    \
            kicks[2 * (uint64_t)sid[q]] = kx[q];                                                              \
            kicks[2 * (uint64_t)sid[q] + 1] = ky[q];                                                          \
        }                                                                                                     \
    done:                                                                                                     \
        free(start); free(fillp); free(key); free(ss); free(sid); free(sx); free(sy); free(kx); free(ky);    \
        free(nx); free(ny); free(ni); free(nkx); free(nky); free(rx); free(ry);                               \
        return pe;                                                                                            \
    }

DEFINE_FAST2(fast_pairs2_f32, float, consts_f, sep_near_f32, LIT_F)
DEFINE_FAST2(fast_pairs2_f64, double, consts_d, sep_near_f64, LIT_D)

static int forces2_full(int prec, const orc_system2 *s, const orc_state2 *st, uint64_t n, const int64_t *pos, const uint32_t *sp, int32_t *kicks, double *pe_out,
                        double *wall_force, double *piston_force, int brute) {
    if (s->nspecies < 1 || s->nspecies > 4) return -3;
    if (prec != 0 && prec != 32 && prec != 64) return -4;
    consts_f cf32[16];
    consts_d cf64[16];
    for (int k = 0; k < 16; k++) {
        cf32[k] = consts_f32(&s->pairs[k]);
        cf64[k] = consts_f64(&s->pairs[k]);
    }
    int32_t *tab = NULL;
    uint32_t rc2[16];
    uint64_t len = 0;
    if (prec == 0) {
        len = orc2_int_table_len(s);
        tab = malloc(s->nspecies * s->nspecies * len * sizeof(int32_t));
        if (!tab) return -2;
        orc2_int_tables(s, tab, rc2);
    }
    uint32_t *partners = partner_table2(s, n);
    uint32_t nc[2];
    present_cells2(s, n, sp, nc);
    if (nc[0] < 3 || nc[1] < 3) {
        /* fewer than 3 cells: md-ref asserts; every pair is a candidate instead */
        brute = 1;
        nc[0] = nc[1] = 1;
    }
    uint64_t ncell = (uint64_t)nc[0] * nc[1];
    uint64_t *start = calloc(ncell + 1, sizeof(uint64_t));
    uint64_t *fill = malloc((ncell + 1) * sizeof(uint64_t));
    uint64_t *items = malloc((n ? n : 1) * sizeof(uint64_t));
    uint32_t *cell = malloc((n ? n : 1) * 2 * sizeof(uint32_t));
    if (!partners || !start || !fill || !items || !cell) return -2;
    double b0 = orc_to_f64(s->box_len[0]), b1 = orc_to_f64(s->box_len[1]);
    for (uint64_t i = 0; i < n; i++) {
        cell[2 * i] = cell_of(orc_to_f64(pos[2 * i]), b0, nc[0]);
        cell[2 * i + 1] = cell_of(orc_to_f64(pos[2 * i + 1]), b1, nc[1]);
        start[(uint64_t)cell[2 * i + 1] * nc[0] + cell[2 * i] + 1] += 1;
    }
    for (uint64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    memcpy(fill, start, (ncell + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) items[fill[(uint64_t)cell[2 * i + 1] * nc[0] + cell[2 * i]]++] = i;
    double pe = 0.0;
    int64_t piston_fixed = 0;
    /* md-ref's fast pair path where IntEngine takes it (int_engine.rs:136-141): the pair kicks
     * first, then the per-particle terms below */
    const int fast = orc2_fast_path && !brute && prec != 0 && fast_eligible2(s, nc);
    if (fast) {
        int err = 0;
        if (prec == 32) {
            float kc[4];
            for (uint32_t a = 0; a < 4; a++) kc[a] = a < s->nspecies ? (float)kick_coeff2(s, a) : 0.0f;
            pe += (double)fast_pairs2_f32(s, n, pos, sp, cf32, kc, nc, kicks, &err);
        } else {
            double kc[4];
            for (uint32_t a = 0; a < 4; a++) kc[a] = a < s->nspecies ? kick_coeff2(s, a) : 0.0;
            pe += fast_pairs2_f64(s, n, pos, sp, cf64, kc, nc, kicks, &err);
        }
        if (err) return err;
    }
    for (uint64_t i = 0; i < n; i++) {
        int32_t k[2] = {0, 0};
        if (fast) {
            k[0] = kicks[2 * i];
            k[1] = kicks[2 * i + 1];
        }
        float pe32 = 0.0f;
        double pe64 = 0.0;
        float c32 = (float)kick_coeff2(s, sp[i]);
        double c64 = kick_coeff2(s, sp[i]);
        uint64_t nblock = fast ? 0 : (brute ? 1 : 9);
        for (uint64_t blk = 0; blk < nblock; blk++) {
            uint64_t q0 = 0, q1 = n;
            if (!brute) {
                uint32_t oy = (uint32_t)(blk / 3), ox = (uint32_t)(blk % 3);
                uint32_t y = (cell[2 * i + 1] + nc[1] + oy - 1) % nc[1];
                uint32_t x = (cell[2 * i] + nc[0] + ox - 1) % nc[0];
                uint64_t c = (uint64_t)y * nc[0] + x;
                q0 = start[c];
                q1 = start[c + 1];
            }
            for (uint64_t q = q0; q < q1; q++) {
                uint64_t j = brute ? q : items[q];
                if (j == i) continue;
                if (!interact_across2(s, &pos[2 * i], &pos[2 * j])) continue;
                if (prec == 32) {
                    pair2_f32(s, i, pos, sp, j, cf32, c32, k, &pe32);
                } else if (prec == 64) {
                    pair2_f64(s, i, pos, sp, j, cf64, c64, k, &pe64);
                } else {
                    int64_t dx = separation2(s, pos[2 * i], pos[2 * j], 0), dy = separation2(s, pos[2 * i + 1], pos[2 * j + 1], 1);
                    int32_t kk[2];
                    if (orc2_int_kick(s, tab, rc2, len, sp[i], sp[j], dx, dy, kk)) {
                        k[0] = (int32_t)((uint32_t)k[0] + (uint32_t)kk[0]);
                        k[1] = (int32_t)((uint32_t)k[1] + (uint32_t)kk[1]);
                        double fx = orc_to_f64(dx), fy = orc_to_f64(dy);
                        pe64 += orc_pair_energy(&s->pairs[sp[i] * 4 + sp[j]], fx * fx + fy * fy);
                    }
                }
            }
        }
        double extra = 0.0;
        if (prec == 32) extras2_f32(s, st, i, pos, partners, c32, k, &extra, wall_force, &piston_fixed);
        else extras2_f64(s, st, i, pos, partners, c64, k, &extra, wall_force, &piston_fixed);
        kicks[2 * i] = k[0];
        kicks[2 * i + 1] = k[1];
        pe += 0.5 * (prec == 32 ? (double)pe32 : pe64) + extra;
    }
    if (pe_out) *pe_out = pe;
    if (piston_force) *piston_force = (double)piston_fixed / PISTON_FORCE_SCALE;
    free(start);
    free(fill);
    free(items);
    free(cell);
    free(tab);
    free(partners);
    return 0;
}

int orc2_forces(int prec, const orc_system2 *sys, const orc_state2 *st, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks, double *pe, double *wall_force, double *piston_force) {
    return forces2_full(prec, sys, st, n, pos, species, kicks, pe, wall_force, piston_force, 0);
}

int orc2_kicks(int prec, const orc_system2 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks, double *pe) {
    orc_state2 *st = malloc(sizeof *st);
    if (!st) return -2;
    orc2_state_init(sys, st);
    int rc = forces2_full(prec, sys, st, n, pos, species, kicks, pe, NULL, NULL, 0);
    free(st);
    return rc;
}

int orc2_kicks_brute(int prec, const orc_system2 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks) {
    orc_state2 *st = malloc(sizeof *st);
    if (!st) return -2;
    orc2_state_init(sys, st);
    int rc = forces2_full(prec, sys, st, n, pos, species, kicks, NULL, NULL, NULL, 1);
    free(st);
    return rc;
}

/* walls::thermal_velocity (walls.rs:100-112), 2D */
static void thermal_velocity2(const orc_system2 *s, uint32_t id, uint32_t clock, double t, double m, double inward, double *vn, double *vt) {
    uint32_t k0 = (uint32_t)s->seed, k1 = (uint32_t)(s->seed >> 32) ^ STREAM_WALL;
    uint32_t r[2], r2[2];
    threefry(id, clock, k0, k1, 20, r);
    threefry(id, clock, k0, k1 ^ 1u, 20, r2);
    double u1 = 1.0 - orc_to_unit_f64(r[0], r[1]);
    double u2 = 1.0 - orc_to_unit_f64(r2[0], 0), u3 = orc_to_unit_f64(r2[1], 0);
    double sd = sqrt(t / m);
    *vn = inward * sd * sqrt(-2.0 * log(u1));
    *vt = sd * sqrt(-2.0 * log(u2)) * cos(TAU * u3);
}

/* walls::andersen (walls.rs:122-139) with Thermostat::acts_on (system.rs:172-176) */
static int andersen2(const orc_system2 *s, uint32_t id, uint32_t clock, uint32_t sp, const int64_t *p, double out[2]) {
    if (!s->has_bath) return 0;
    if (!((s->bath_species >> sp) & 1u)) return 0;
    int64_t x = p[s->bath_axis];
    if (!(x >= s->bath_region[0] && x < s->bath_region[1])) return 0;
    double m = s->masses[sp];
    uint32_t k0 = (uint32_t)s->seed, k1 = (uint32_t)(s->seed >> 32) ^ STREAM_THERMO;
    uint32_t r[2], r2[2];
    threefry(id, clock, k0, k1, 20, r);
    if (orc_to_unit_f64(r[0], 0) >= s->bath_rate * s->dt) return 0;
    threefry(id, clock, k0, k1 ^ 1u, 20, r2);
    double u1 = 1.0 - orc_to_unit_f64(r[1], 0);
    double u2 = orc_to_unit_f64(r2[0], 0);
    double sd = sqrt(s->bath_t / m);
    double rad = sqrt(-2.0 * log(u1));
    double a = TAU * u2;
    out[0] = sd * rad * cos(a);
    out[1] = sd * rad * sin(a);
    return 1;
}

static void book2(orc_state2 *st, int w, uint64_t c, double de, double dp, int hit) {
    st->ledger.energy[w] += de;
    st->ledger.impulse[w] += dp;
    st->ledger.hits[w] += (uint64_t)hit;
    if (c < ORC_MAX_COMPARTMENTS) {
        st->comp_energy[w][c] += de;
        st->comp_impulse[w][c] += dp;
        st->comp_hits[w][c] += (uint64_t)hit;
    }
}

/* walls::piston_kick (walls.rs:143-146) */
static int32_t piston_kick2(const orc_system2 *s, double net, double mass) { return sat_i32(round(net * s->dt * s->dt / mass * ONE_F)); }

/* IntEngine::step (int_engine.rs:253-344) with move_piston (:346-361) and collide (:364-387) */
int orc2_run(int prec, const orc_system2 *s, orc_state2 *st, uint64_t n, int64_t *pos, int32_t *vel, const uint32_t *species, uint64_t steps) {
    int32_t *kicks = malloc((n ? n : 1) * 2 * sizeof(int32_t));
    double *wf = malloc((n ? n : 1) * WALLS2 * sizeof(double));
    if (!kicks || !wf) return -2;
    int32_t g[2] = {gravity_kick(s->gravity[0], s->dt), gravity_kick(s->gravity[1], s->dt)};
    int rc = 0;
    for (uint64_t stp = 0; stp < steps; stp++) {
        memset(wf, 0, n * WALLS2 * sizeof(double));
        double piston_force = 0.0;
        rc = forces2_full(prec, s, st, n, pos, species, kicks, NULL, wf, &piston_force, 0);
        if (rc) break;
        for (uint64_t i = 0; i < n; i++)
            for (int w = 0; w < WALLS2; w++)
                if (wf[WALLS2 * i + w] != 0.0) {
                    int axis = w / 2;
                    uint64_t c = (uint64_t)compartment_of2(s, pos[2 * i + (1 - axis)], 1 - axis);
                    book2(st, w, c, 0.0, wf[WALLS2 * i + w] * s->dt, 0);
                }
        for (uint64_t i = 0; i < n; i++) {
            uint32_t sp = species[i];
            int pin = pinned2(s, sp);
            double m = s->masses[sp];
            int32_t v[2];
            int64_t x[2] = {pos[2 * i], pos[2 * i + 1]};
            int64_t home[2] = {compartment_of2(s, x[0], 0), compartment_of2(s, x[1], 1)};
            for (int a = 0; a < 2; a++) v[a] = (int32_t)((uint32_t)vel[2 * i + a] + (uint32_t)kicks[2 * i + a] + (uint32_t)(pin ? 0 : g[a]));
            double u[2];
            if (!pin && andersen2(s, (uint32_t)i, st->clock, sp, x, u)) {
                double before[2] = {velocity2(s, v[0]), velocity2(s, v[1])};
                v[0] = to_step2(s, u[0]);
                v[1] = to_step2(s, u[1]);
                v[0] = (int32_t)((uint32_t)v[0] + (uint32_t)to_step2(s, s->bath_drift));
                if (s->has_piston && !s->piston_free) {
                    double f = orc_to_f64(x[1]) / orc_to_f64(st->piston_w);
                    f = fmin(fmax(f, 0.0), 1.0);
                    v[1] = (int32_t)((uint32_t)v[1] + (uint32_t)sat_i32(round((double)st->piston_v * f)));
                }
                double after[2] = {velocity2(s, v[0]), velocity2(s, v[1])};
                double de = 0.5 * m * (after[0] * after[0] + after[1] * after[1] - before[0] * before[0] - before[1] * before[1]);
                book2(st, BATH2, 0, de, 0.0, 1);
            }
            int64_t old[2] = {x[0], x[1]};
            for (int a = 0; a < 2; a++) x[a] += (int64_t)v[a];
            if (s->has_membrane) {
                int a = (int)s->membrane_axis;
                int64_t along = old[1 - a];
                int in_pore = s->pore_shift != 0 && (along & (((int64_t)1 << s->pore_shift) - 1)) < s->pore_width;
                if (((s->membrane_blocks >> sp) & 1u) && !in_pore) {
                    int64_t w = s->membrane_position;
                    int lo_old = old[a] < w, lo_new = x[a] < w;
                    if (lo_old != lo_new) {
                        int face = lo_old ? MEMBRANE_LOW : MEMBRANE_HIGH;
                        x[a] = 2 * w - 1 - x[a];
                        double dp = 2.0 * m * fabs(velocity2(s, v[a]));
                        v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
                        book2(st, face, (uint64_t)compartment_of2(s, pos[2 * i + (1 - a)], 1 - a), 0.0, dp, 1);
                    }
                }
            }
            for (int a = 0; a < 2; a++) {
                int64_t l = s->box_len[a];
                if (!s->walled[a]) {
                    x[a] = orc_wrap(x[a], l);
                    continue;
                }
                int side = -1;
                if (x[a] < 0) {
                    x[a] = -1 - x[a];
                    side = 0;
                } else if (x[a] >= l) {
                    x[a] = 2 * l - 1 - x[a];
                    side = 1;
                }
                if (side >= 0) {
                    int w = 2 * a + side;
                    double before[2] = {velocity2(s, v[0]), velocity2(s, v[1])};
                    if (s->wall_kind[a][side] == ORC_WALL_THERMAL) {
                        double vn, vt;
                        thermal_velocity2(s, (uint32_t)i, st->clock, s->wall_t[a][side], m, side == 0 ? 1.0 : -1.0, &vn, &vt);
                        v[a] = to_step2(s, vn);
                        v[1 - a] = to_step2(s, vt);
                    } else {
                        v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
                    }
                    double after[2] = {velocity2(s, v[0]), velocity2(s, v[1])};
                    double kea = 0.5 * m * (after[0] * after[0] + after[1] * after[1]);
                    double keb = 0.5 * m * (before[0] * before[0] + before[1] * before[1]);
                    book2(st, w, (uint64_t)compartment_of2(s, pos[2 * i + (1 - a)], 1 - a), kea - keb, fabs(m * (before[a] - after[a])), 1);
                } else if (s->compartments[a] != 0) {
                    uint32_t shift = s->compartments[a];
                    if ((x[a] >> shift) != home[a]) {
                        int64_t lo = home[a] << shift, hi = (home[a] + 1) << shift;
                        x[a] = x[a] < lo ? 2 * lo - 1 - x[a] : (x[a] >= hi ? 2 * hi - 1 - x[a] : x[a]);
                        v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
                    }
                }
            }
            if (s->has_piston && x[1] >= st->piston_w) {
                x[1] = 2 * st->piston_w - 1 - x[1];
                v[1] = (int32_t)((uint32_t)0 - (uint32_t)v[1]);
            }
            vel[2 * i] = v[0];
            vel[2 * i + 1] = v[1];
            pos[2 * i] = x[0];
            pos[2 * i + 1] = x[1];
        }
        if (s->has_piston) {
            if (s->piston_free) st->piston_v = (int32_t)((uint32_t)st->piston_v + (uint32_t)piston_kick2(s, piston_force - s->piston_force, s->piston_mass));
            int64_t old = st->piston_w;
            st->piston_w += (int64_t)st->piston_v;
            int64_t lo = orc_to_fixed(s->piston_min), hi = orc_to_fixed(s->piston_max);
            if (st->piston_w < lo || st->piston_w > hi) {
                st->piston_w = st->piston_w < lo ? lo : hi;
                st->piston_v = 0;
            }
            st->piston_work += piston_force * orc_to_f64(st->piston_w - old);
        }
        st->clock += 1;
    }
    free(kicks);
    free(wf);
    return rc;
}

int orc2_step(int prec, const orc_system2 *s, uint64_t n, int64_t *pos, int32_t *vel, const uint32_t *species, uint64_t steps) {
    orc_state2 *st = malloc(sizeof *st);
    if (!st) return -2;
    orc2_state_init(s, st);
    int rc = orc2_run(prec, s, st, n, pos, vel, species, steps);
    free(st);
    return rc;
}

/* IntEngine::reverse (int_engine.rs:392-409) */
int orc2_reverse_run(int prec, const orc_system2 *s, orc_state2 *st, uint64_t n, const int64_t *pos, int32_t *vel, const uint32_t *species) {
    int32_t *kicks = malloc((n ? n : 1) * 2 * sizeof(int32_t));
    if (!kicks) return -2;
    double piston_force = 0.0;
    int rc = forces2_full(prec, s, st, n, pos, species, kicks, NULL, NULL, &piston_force, 0);
    if (!rc) {
        for (uint64_t i = 0; i < n; i++) {
            int pin = pinned2(s, species[i]);
            for (int a = 0; a < 2; a++) {
                uint32_t g = pin ? 0 : (uint32_t)gravity_kick(s->gravity[a], s->dt);
                vel[2 * i + a] = (int32_t)((uint32_t)0 - ((uint32_t)vel[2 * i + a] + (uint32_t)kicks[2 * i + a] + g));
            }
        }
        if (s->has_piston) {
            int32_t kick = s->piston_free ? piston_kick2(s, piston_force - s->piston_force, s->piston_mass) : 0;
            st->piston_v = (int32_t)((uint32_t)0 - ((uint32_t)st->piston_v + (uint32_t)kick));
        }
    }
    free(kicks);
    return rc;
}

int orc2_reverse(int prec, const orc_system2 *s, uint64_t n, const int64_t *pos, int32_t *vel, const uint32_t *species) {
    orc_state2 *st = malloc(sizeof *st);
    if (!st) return -2;
    orc2_state_init(s, st);
    int rc = orc2_reverse_run(prec, s, st, n, pos, vel, species);
    free(st);
    return rc;
}

/* IntEngine::energy (int_engine.rs:417-446) */
int orc2_energy(int prec, const orc_system2 *s, uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species, double *kinetic, double *potential) {
    int32_t *kicks = malloc((n ? n : 1) * 2 * sizeof(int32_t));
    orc_state2 *st = malloc(sizeof *st);
    if (!kicks || !st) return -2;
    orc2_state_init(s, st);
    double pe = 0.0, ke = 0.0;
    int rc = forces2_full(prec, s, st, n, pos, species, kicks, &pe, NULL, NULL, 0);
    if (!rc)
        for (uint64_t i = 0; i < n; i++) {
            double m = s->masses[species[i]];
            double v[2];
            for (int a = 0; a < 2; a++) {
                int32_t next = (int32_t)((uint32_t)vel[2 * i + a] + (uint32_t)kicks[2 * i + a] + (uint32_t)gravity_kick(s->gravity[a], s->dt));
                v[a] = 0.5 * (velocity2(s, vel[2 * i + a]) + velocity2(s, next));
            }
            ke += 0.5 * m * (v[0] * v[0] + v[1] * v[1]);
            for (int a = 0; a < 2; a++) pe -= m * s->gravity[a] * orc_to_f64(pos[2 * i + a]);
        }
    *kinetic = ke;
    *potential = pe;
    free(kicks);
    free(st);
    return rc;
}

/* crates/md/src/checks.rs:156-164 */
void orc2_golden_setup(orc_system2 *sys, int64_t *pos, int32_t *vel, uint32_t *species) {
    memset(sys, 0, sizeof *sys);
    orc2_box_for(3000, 0.6, 1.0, sys->box_len);
    sys->dt = 0.005;
    sys->nspecies = 2;
    sys->masses[0] = 1.0;
    sys->masses[1] = 2.0;
    orc_pair lj = {1.0, 1.0, 2.5}, cross = {0.7, 1.05, 2.5};
    for (int k = 0; k < 16; k++) sys->pairs[k] = lj;
    sys->pairs[0 * 4 + 1] = cross;
    sys->pairs[1 * 4 + 0] = cross;
    sys->seed = 1;
    orc2_fluid(sys, 3000, 1.0, 2026, pos, vel, species);
    for (int i = 0; i < 3000; i += 4) species[i] = 1;
    orc2_maxwell(sys, 3000, species, 1.0, 2026, vel);
}
