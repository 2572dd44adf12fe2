// md-core, the engine's data contract, on the host: fixed-point units, the LJ pair, seeded
// initial conditions (Threefry streams and the `libm` crate's Box-Muller) and the state hash.
// Mirrors crates/md-core/src/{fixed,pair,init,three,system}.rs of the reference.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/rahman_hip.h"

namespace rh::core {

constexpr int64_t ONE = int64_t(1) << 32;
constexpr double ONE_F = 4294967296.0;
constexpr double UNIT = 1.0 / 4294967296.0;

// Rust `as` conversions from f64: truncation toward zero, saturating, NaN -> 0.
int32_t sat_i32(double x);
int64_t sat_i64(double x);
uint32_t sat_u32(double x);
uint64_t sat_u64(double x);

int64_t to_fixed(double x);  // fixed.rs:21-23
inline double to_f64(int64_t x) { return (double)x * UNIT; }
inline int64_t wrap(int64_t x, int64_t l) { return x < 0 ? x + l : (x >= l ? x - l : x); }

// x^n for the constant exponents the reference uses, by LLVM's binary powering.
double powi(double x, int n);

double pair_shift(const rh_pair& p);  // pair.rs:29-35
double pair_f_over_r(const rh_pair& p, double r2);
double pair_energy(const rh_pair& p, double r2);

// Threefry streams (gpu-prims threefry.rs:29-46) and the FreeBSD-derived log/cos/sin the
// reference's `libm` crate provides.
double to_unit_f64(uint32_t r0, uint32_t r1);
double uniform(uint64_t seed, uint32_t stream, uint32_t i);
double normal(uint64_t seed, uint32_t stream, uint32_t i);
double libm_log(double x);
double libm_cos(double x);
double libm_sin(double x);

double rc_max(const rh_pair* pairs, uint32_t ns);
double kick_coeff(double dt, double mass, bool pinned);
int32_t to_step_displacement(double dt, double v);
int32_t gravity_kick(double dt, double g);

// md_core::table::IntTables (table.rs:17-73): all-integer pair forces for the 2D engine.  Entry
// k of species pair (a, b) is round(F/r(r2) c_a 2^T_BITS) at r2 = k 2^STEP_BITS / 2^R2_BITS
// (clamped below r2 = 0.36), continued one step past the cutoff; rc2 in 2^-R2_BITS sigma^2.
constexpr int TAB_SEP_BITS = 20, TAB_R2_BITS = 24, TAB_T_BITS = 2, TAB_STEP_BITS = 12;
constexpr int TAB_KICK_SHIFT = TAB_T_BITS + TAB_SEP_BITS;
struct IntTables {
    uint32_t ns = 0;
    uint64_t len = 0;
    std::vector<int32_t> data;  // [(a * ns + b) * len + k]
    uint32_t rc2[16] = {};      // [a * ns + b]
};
IntTables int_tables(const rh_pair* pairs, uint32_t ns, const double* masses, uint32_t pinned, double dt);

}  // namespace rh::core
