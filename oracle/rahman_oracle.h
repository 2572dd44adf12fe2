/*
 * rahman_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in plain C, of the reference engine's CPU twin (md-ref) and of the parts of
 * its shared model (md-core, gpu-prims' Threefry) that the twin needs.  It is the checker the
 * HIP engine is compared against.  Only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark load it; the product library never links or calls it.
 *
 * Every function names the reference file:line it follows (paths relative to the reference
 * repository root).  Numeric conventions (from the Rust originals):
 *   - float -> int conversions saturate and map NaN to 0 (Rust `as`);
 *   - `round_ties_even` is rint() under the default rounding mode, `f64::round` is round();
 *   - powi(x, n) is evaluated by the binary powering LLVM uses for constant exponents;
 *   - everything is compiled with -ffp-contract=off (no fused multiply-adds).
 *
 * Parity pins (what checks each part; see tests/test_oracle.py):
 *   threefry        Random123 known-answer vectors (crates/gpu-prims/src/threefry.rs:53-70)
 *   libm log/cos    glibc on 10^5 arguments (<= 1 ulp, >= 99% identical) and the golden hash
 *   fixed           identities of crates/md-core/src/fixed.rs:74-99
 *   pair            closed forms of crates/md-core/src/pair.rs:63-83
 *   init (2D)       the all-integer golden hash 0x3aaeff680a2f0278 (crates/md/src/checks.rs:156-167)
 *   init (3D)       fcc spacing / fcc_cells / Maxwell identities (crates/md-core/src/three.rs:410-446);
 *                   cbrt comes from the C library, as Rust's f64::cbrt does on Linux: parity unpinned
 *   int tables (2D) the golden hash, plus antisymmetry (crates/md-core/src/table.rs:113-136)
 *   forces (2D/3D)  brute-force O(N^2) kicks on small systems (exact: integer sums), exact
 *                   reversal, zero momentum, permutation invariance, f32-vs-f64 rms < 1e-5
 *                   (crates/md-ref/tests/physics.rs), energy conservation
 */
#ifndef RAHMAN_ORACLE_H
#define RAHMAN_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Test hook: plant a known mistake in the 3D pair loop (see rahman_oracle.c; 0 = none). */
void orc_set_mutation(int m);

/* ---- gpu-prims: Threefry-2x32 and its counter-based streams ---- */
void orc_threefry2x32_rounds(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t rounds, uint32_t out[2]);
double orc_to_unit_f64(uint32_t r0, uint32_t r1);
double orc_uniform(uint64_t seed, uint32_t stream, uint32_t i);
double orc_normal(uint64_t seed, uint32_t stream, uint32_t i);

/* ---- the libm routines the reference's `libm` crate supplies ---- */
double orc_log(double x);
double orc_cos(double x);
double orc_sin(double x);

/* ---- md-core fixed point ---- */
int64_t orc_to_fixed(double x);
double orc_to_f64(int64_t x);
int64_t orc_min_image(int64_t a, int64_t b, int64_t l);
int64_t orc_wrap(int64_t x, int64_t l);
/* 1 and *out set if |d| < 8 sigma, else 0. */
int orc_sep_f32(int64_t d, float *out);

/* ---- md-core Pair ---- */
typedef struct {
    double eps, sigma, rc;
} orc_pair;
double orc_pair_shift(const orc_pair *p);
double orc_pair_energy(const orc_pair *p, double r2);
double orc_pair_f_over_r(const orc_pair *p, double r2);

/* ---- md-core System3 (crates/md-core/src/three.rs:13-31) ---- */
enum { ORC_WALL_HARD = 0, ORC_WALL_SOFT = 1, ORC_WALL_THERMAL = 2 };

typedef struct {
    int64_t box_len[3];
    double dt;
    uint32_t nspecies;
    uint32_t pinned;          /* bit s: species s pinned */
    double masses[4];
    orc_pair pairs[16];       /* [a*4 + b], symmetric */
    uint32_t walled[3];       /* 0: periodic axis, 1: walls at 0 and L */
    uint32_t wall_kind[3][2]; /* ORC_WALL_* per axis and side */
    double wall_t[3][2];      /* temperature of thermal walls */
    double gravity[3];
    uint64_t seed;
    uint32_t has_bath;
    uint32_t bath_species;
    double bath_t, bath_rate;
} orc_system3;

typedef struct {
    double energy[7], impulse[7];
    uint64_t hits[7];
} orc_ledger3;

/* init */
void orc3_cube_for(uint64_t n, double density, int64_t out[3]);
void orc3_fcc_cells(uint64_t n, const int64_t box_len[3], uint64_t out[3]);
void orc3_fcc(uint64_t n, const int64_t box_len[3], const uint64_t cells[3], double jitter, uint64_t seed, int64_t *pos);
void orc3_maxwell(const orc_system3 *sys, uint64_t n, const uint32_t *species, double t, uint64_t seed, int32_t *vel);
void orc3_zero_velocity_sum(const orc_system3 *sys, uint64_t n, const uint32_t *species, int32_t *vel);
void orc3_set_temperature(const orc_system3 *sys, uint64_t n, const uint32_t *species, int32_t *vel, double t);
uint64_t orc3_hash(uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species);
void orc3_cells_for(const orc_system3 *sys, double r, uint32_t out[3]);

/* IntEngine3<F> (crates/md-ref/src/int_engine3.rs); `prec` is 32 (f32) or 64 (f64).
 * Arrays are AoS: pos[3i+a], vel[3i+a], kicks[3i+a].  Return 0 on success. */
int orc3_kicks(int prec, const orc_system3 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks, double *pe);
/* brute force over all pairs (no cell list), for pinning the cell list */
int orc3_kicks_brute(int prec, const orc_system3 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks);
int orc3_step(int prec, const orc_system3 *sys, uint64_t n, int64_t *pos, int32_t *vel, const uint32_t *species, uint64_t steps, uint32_t *clock, orc_ledger3 *ledger);
int orc3_reverse(int prec, const orc_system3 *sys, uint64_t n, const int64_t *pos, int32_t *vel, const uint32_t *species);
/* IntEngine3::energy (on-step kinetic + potential), always evaluated with F = f64 when prec = 64 */
int orc3_energy(int prec, const orc_system3 *sys, uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species, double *kinetic, double *potential);

/* ---- md-core System (2D, periodic / hard-walled subset) ---- */
typedef struct {
    int64_t box_len[2];
    double dt;
    uint32_t nspecies;
    uint32_t pinned;
    double masses[4];
    orc_pair pairs[16];
    uint32_t walled[2]; /* walls on this axis (kinds below; all zero: hard) */
    double gravity[2];
    uint64_t seed;
    /* md_core::System's boundaries (crates/md-core/src/system.rs:26-61, 71-180) */
    uint32_t wall_kind[2][2];   /* ORC_WALL_* per axis and side */
    double wall_t[2][2];
    uint32_t compartments[2];   /* shift (log2 of the side in position units), 0: none */
    uint32_t diathermal;        /* CompartmentWalls::Diathermal */
    uint32_t has_bath, bath_species, bath_axis;
    double bath_t, bath_rate, bath_drift;
    int64_t bath_region[2];
    uint32_t has_membrane, membrane_axis, membrane_blocks, pore_shift;
    int64_t membrane_position, pore_width;
    uint32_t has_piston, piston_free;
    double piston_speed, piston_mass, piston_force, piston_start, piston_min, piston_max;
    uint32_t nbond_types;
    uint32_t bond_fene[4];
    double bond_k[4], bond_r0[4];
    uint64_t nbonds;
    const uint32_t *bonds;      /* [id_a, id_b, type] per bond */
} orc_system2;

/* IntEngine's run state beyond positions and velocities (int_engine.rs:23-48): the clock of the
 * stochastic boundaries, the piston, and the ledger in total and per compartment along each
 * wall. */
#define ORC_MAX_COMPARTMENTS 256
typedef struct {
    uint32_t clock;
    int64_t piston_w;
    int32_t piston_v;
    double piston_work;
    orc_ledger3 ledger;
    double comp_energy[7][ORC_MAX_COMPARTMENTS], comp_impulse[7][ORC_MAX_COMPARTMENTS];
    uint64_t comp_hits[7][ORC_MAX_COMPARTMENTS];
} orc_state2;

void orc2_state_init(const orc_system2 *sys, orc_state2 *st);
/* IntEngine::forces (int_engine.rs:130-251): kicks, potential, per-particle force on each wall
 * (wall_force[7 i + w], may be NULL) and the piston force (may be NULL). */
int orc2_forces(int prec, const orc_system2 *sys, const orc_state2 *st, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks, double *pe, double *wall_force, double *piston_force);
/* IntEngine::step / reverse with the full 2D boundaries (int_engine.rs:253-400) */
int orc2_run(int prec, const orc_system2 *sys, orc_state2 *st, uint64_t n, int64_t *pos, int32_t *vel, const uint32_t *species, uint64_t steps);
int orc2_reverse_run(int prec, const orc_system2 *sys, orc_state2 *st, uint64_t n, const int64_t *pos, int32_t *vel, const uint32_t *species);

void orc2_box_for(uint64_t n, double density, double aspect, int64_t out[2]);
void orc2_lattice(uint64_t n, const int64_t box_len[2], double jitter, uint64_t seed, int64_t *pos);
void orc2_maxwell(const orc_system2 *sys, uint64_t n, const uint32_t *species, double t, uint64_t seed, int32_t *vel);
void orc2_fluid(const orc_system2 *sys, uint64_t n, double t, uint64_t seed, int64_t *pos, int32_t *vel, uint32_t *species);
uint64_t orc2_hash(uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species);
void orc2_cells_for(const orc_system2 *sys, double r, uint32_t out[2]);

/* prec: 32, 64, or 0 for the all-integer table forces (IntEngine::integer). */
int orc2_kicks(int prec, const orc_system2 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks, double *pe);
int orc2_kicks_brute(int prec, const orc_system2 *sys, uint64_t n, const int64_t *pos, const uint32_t *species, int32_t *kicks);
int orc2_step(int prec, const orc_system2 *sys, uint64_t n, int64_t *pos, int32_t *vel, const uint32_t *species, uint64_t steps);
/* md-ref's IntEngine takes its fast pair path (crates/md-ref/src/fast.rs) by default; the oracle
 * runs the reference loop unless this is set (0 / 1). */
void orc2_set_fast_path(int on);
/* md_core::three::andersen and thermal_velocity (three.rs:232-265) with the C library's
 * log/cos/sin, every key redrawn: 6 integer displacements per key (bath x3, wall normal and
 * tangentials). */
void orc3_draws(uint64_t seed, double t, double mass, double dt, uint32_t id0, uint64_t count, uint32_t clock, int32_t *out);
int orc2_reverse(int prec, const orc_system2 *sys, uint64_t n, const int64_t *pos, int32_t *vel, const uint32_t *species);
int orc2_energy(int prec, const orc_system2 *sys, uint64_t n, const int64_t *pos, const int32_t *vel, const uint32_t *species, double *kinetic, double *potential);

/* md-core IntTables (crates/md-core/src/table.rs:31-106) */
uint64_t orc2_int_table_len(const orc_system2 *sys);
void orc2_int_tables(const orc_system2 *sys, int32_t *data, uint32_t *rc2);
int orc2_int_kick(const orc_system2 *sys, const int32_t *data, const uint32_t *rc2, uint64_t len, uint32_t sa, uint32_t sb, int64_t dx, int64_t dy, int32_t out[2]);

/* The golden all-integer setup (crates/md/src/checks.rs:156-164): fills sys and the state of 3000. */
void orc2_golden_setup(orc_system2 *sys, int64_t *pos, int32_t *vel, uint32_t *species);

#ifdef __cplusplus
}
#endif
#endif
