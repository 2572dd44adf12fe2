/*
 * rahman_hip.h -- C ABI of the MI355X-native Rahman molecular-dynamics engine.
 *
 * The boundary mirrors the reference's Rust engine API (kshitijl/micrographia):
 *   - rh_system3 / rh_system2 are POD images of md_core::three::System3 and md_core::System
 *     (crates/md-core/src/three.rs:13-31, crates/md-core/src/system.rs:26-61);
 *   - a state is passed as three flat arrays, positions i64[D*n], velocities i32[D*n] and
 *     species u32[n], in particle-id order: md_core::three::State3 (three.rs:267-273);
 *   - rh_engine3_* replaces md::GpuEngine3 (crates/md/src/engine3.rs:307-648) and rh_engine2_*
 *     md::GpuEngine (crates/md/src/engine.rs:567-1098).
 * Every call returns RH_OK or a negative status; rh_last_error() describes the last failure on
 * the calling thread.  Nothing unwinds or aborts across the boundary.  An engine handle belongs
 * to one host thread; it owns one HIP stream; step() returns once the steps are done on the
 * device, reads synchronise.
 */
#ifndef RAHMAN_HIP_H
#define RAHMAN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RH_OK 0
#define RH_ERR_INVALID (-1)   /* a contract violation (the reference's assert!) */
#define RH_ERR_HIP (-2)       /* a HIP runtime failure */
#define RH_ERR_NOMEM (-3)     /* allocation failed or over budget */
#define RH_ERR_UNSUPPORTED (-4)

#define RH_MAX_SPECIES 4

/* md_core::Pair (crates/md-core/src/pair.rs:3-8) */
typedef struct rh_pair {
    double eps, sigma, rc;
} rh_pair;

/* md_core::Wall (crates/md-core/src/system.rs:12-24) */
enum { RH_WALL_HARD = 0, RH_WALL_SOFT = 1, RH_WALL_THERMAL = 2 };

/* md_core::three::System3 (crates/md-core/src/three.rs:13-31).  pairs are indexed [a*4 + b]
 * (symmetric); walled[a] = 0 for a periodic axis; bath fields are the Andersen thermostat. */
typedef struct rh_system3 {
    int64_t box_len[3];
    double dt;
    uint32_t nspecies;
    uint32_t pinned;
    double masses[4];
    rh_pair pairs[16];
    uint32_t walled[3];
    uint32_t wall_kind[3][2];
    double wall_t[3][2];
    double gravity[3];
    uint64_t seed;
    uint32_t has_bath;
    uint32_t bath_species;
    double bath_t, bath_rate;
} rh_system3;

/* md_core::System (crates/md-core/src/system.rs:26-61) with its boundaries (system.rs:71-180).
 * Zero-initialised boundary fields mean "none" (hard walls on walled axes, no compartments,
 * thermostat, membrane, piston or bonds).
 *   wall_kind / wall_t   Wall per walled axis and side (RH_WALL_*), temperature of heat-bath walls
 *   compartments[a]      0, or log2 of the compartment side in position units (>= 32): hard walls
 *                        every 2^shift along a walled axis (System::with_compartments)
 *   diathermal           CompartmentWalls::Diathermal (pairs interact across compartment walls)
 *   has_bath ...         Thermostat {t, rate, species, region [lo, hi) along bath_axis (position
 *                        units), drift (sigma/tau, along x)}; a caller without a region passes
 *                        INT64_MIN, INT64_MAX (System::with_thermostat's default)
 *   has_membrane ...     Membrane {axis, position (position units), blocks (bit per species),
 *                        pore_shift (0: no pores), pore_width (position units)}
 *   has_piston ...       Piston {drive: piston_free = 0 Speed(piston_speed), 1 Free {piston_mass,
 *                        piston_force}; start, min, max (sigma)}
 *   bond_* / bonds       BondType per type (bond_fene[t] = 1: Fene, 0: Harmonic; k, r0) and
 *                        nbonds triples [id_a, id_b, type] (the caller keeps them alive during
 *                        the call that takes the system). */
typedef struct rh_system2 {
    int64_t box_len[2];
    double dt;
    uint32_t nspecies;
    uint32_t pinned;
    double masses[4];
    rh_pair pairs[16];
    uint32_t walled[2];
    double gravity[2];
    uint64_t seed;
    uint32_t wall_kind[2][2];
    double wall_t[2][2];
    uint32_t compartments[2];
    uint32_t diathermal;
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
    const uint32_t* bonds;
} rh_system2;

#define RH_MAX_COMPARTMENTS 256 /* md_core::system::MAX_COMPARTMENTS (system.rs:178) */
#define RH_MAX_BONDS 4          /* md_core::system::MAX_BONDS (system.rs:67) */

/* md::ForceMode (crates/md/src/engine.rs:104-110): kind 0 = CellList, 1 = Verlet{skin, cap}.
 * cap 0 asks for the engine's own capacity rule. */
typedef struct rh_mode {
    uint32_t kind;
    uint32_t cap;
    double skin;
} rh_mode;

/* md::engine3::Ledger3 (crates/md/src/engine3.rs:199-205) */
typedef struct rh_ledger3 {
    double energy[7], impulse[7];
    uint64_t hits[7];
} rh_ledger3;

/* md_core::walls::Ledger (crates/md-core/src/walls.rs:10-13, 41-57): entries 0-3 the walls (2 axis +
 * side), 4 and 5 the membrane's low and high faces, 6 the thermostat. */
typedef struct rh_ledger2 {
    double energy[7], impulse[7];
    uint64_t hits[7];
} rh_ledger2;

/* Counters of the list machinery (the reference's read_list_stats, engine3.rs:272-279, plus
 * this engine's own). */
typedef struct rh_list_stats {
    uint32_t rebuilds;
    uint32_t overflowed;  /* a list ever overflowed (it was then regrown and rebuilt) */
    uint32_t cap;         /* current capacity per particle */
    uint32_t max_count;   /* largest list at the last build */
    uint64_t steps_done;
    uint64_t launches_skipped; /* gated step launches that found the list stale */
} rh_list_stats;

const char* rh_last_error(void);
const char* rh_version(void);
int rh_device_count(int* count);

/* ---- md-core initial conditions (crates/md-core/src/three.rs:311-404, init.rs:15-121) ---- */
int rh_cube_for(uint64_t n, double density, int64_t out[3]);
int rh_fcc_cells(uint64_t n, const int64_t box_len[3], uint64_t out[3]);
int rh_fcc(uint64_t n, const int64_t box_len[3], const uint64_t cells[3], double jitter, uint64_t seed, int64_t* pos);
int rh_maxwell3(const rh_system3* sys, uint64_t n, const uint32_t* species, double t, uint64_t seed, int32_t* vel);
int rh_set_temperature3(const rh_system3* sys, uint64_t n, const uint32_t* species, int32_t* vel, double t);
uint64_t rh_hash3(uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species);
int rh_box_for(uint64_t n, double density, double aspect, int64_t out[2]);
int rh_lattice(uint64_t n, const int64_t box_len[2], double jitter, uint64_t seed, int64_t* pos);
int rh_maxwell2(const rh_system2* sys, uint64_t n, const uint32_t* species, double t, uint64_t seed, int32_t* vel);
int rh_fluid2(const rh_system2* sys, uint64_t n, double t, uint64_t seed, int64_t* pos, int32_t* vel, uint32_t* species);
uint64_t rh_hash2(uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species);
/* ForceMode::verlet_with_skin (engine3.rs:61-68 / engine.rs:129-141) without the 128 MiB
 * WebGPU binding clamp. */
int rh_verlet_mode3(const rh_system3* sys, uint64_t n, double skin, rh_mode* out);
int rh_verlet_mode2(const rh_system2* sys, uint64_t n, double skin, rh_mode* out);

/* ---- the 3D engine: md::GpuEngine3 ---- */
typedef struct rh_engine3 rh_engine3;

/* GpuEngine3::with_mode (engine3.rs:342-457); mode NULL = GpuEngine3::new (Verlet, skin 0.4). */
int rh_engine3_create(const rh_system3* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, const rh_mode* mode, rh_engine3** out);
int rh_engine3_destroy(rh_engine3* e);
/* Device memory (bytes) an engine for this system would allocate, computed on the host without
 * allocating (no reference counterpart: the serving loop's admission budget), and the memory an
 * engine holds now (the same figure, plus any list regrowth). */
int rh_engine3_bytes_needed(const rh_system3* sys, uint64_t n, const rh_mode* mode, uint64_t* bytes);
int rh_engine3_device_bytes(const rh_engine3* e, uint64_t* bytes);
int rh_engine3_n(const rh_engine3* e, uint64_t* n);
int rh_engine3_steps(const rh_engine3* e, int64_t* steps);
int rh_engine3_step(rh_engine3* e, uint64_t steps);                     /* engine3.rs:491-522 */
int rh_engine3_reverse(rh_engine3* e);                                  /* engine3.rs:524-531 */
int rh_engine3_update_system(rh_engine3* e, const rh_system3* sys);     /* engine3.rs:533-559 */
int rh_engine3_kick_region(rh_engine3* e, const double lo[3], const double hi[3], double scale, const double add[3]); /* :561-568 */
int rh_engine3_set_species(rh_engine3* e, const uint32_t* ids, uint64_t count, uint32_t species);             /* :570-581 */
int rh_engine3_set_species_velocity(rh_engine3* e, uint32_t species, const double v[3]);                      /* :583-589 */
int rh_engine3_upload(rh_engine3* e, const int64_t* pos, const int32_t* vel, const uint32_t* species);          /* :591-602 */
int rh_engine3_read_state(rh_engine3* e, int64_t* pos, int32_t* vel, uint32_t* species);                        /* :604-606 */
int rh_engine3_kicks(rh_engine3* e, int32_t* kicks, double* potential);                                         /* :608-616 */
int rh_engine3_measure_kinetic(rh_engine3* e);                                                                  /* :618-624 */
/* Kinetic energies (f32; computed in f64 from the half-step velocities, so within one f32
 * rounding of the reference's f32 values) and species, both in particle-id order. */
int rh_engine3_read_kinetic(rh_engine3* e, float* ke, uint32_t* species);                                       /* :626-631 */
int rh_engine3_read_kinetic_total(rh_engine3* e, double* ke);                                                   /* :633-637 */
int rh_engine3_read_potential(rh_engine3* e, double* pe);                                                       /* :640-643 */
int rh_engine3_read_ledger(rh_engine3* e, rh_ledger3* out);                                                     /* :645-647 */
int rh_engine3_list_stats(rh_engine3* e, rh_list_stats* out);                                                   /* :272-279 */
int rh_engine3_set_sort_interval(rh_engine3* e, uint32_t interval);                                             /* sort_interval, :327-329 */
/* Kernel choice (this engine's own tuning knob, no reference counterpart): variant 0 = full
 * Verlet list, one lane per particle; 1 = half list inside work-group blocks of `block` particles
 * (256, 512, 1024, 2048 or 4096; the partner's kick accumulated in LDS) and full lists across blocks.
 * The default is variant 1 with blocks of 512, 1024 from 2^18 particles, 2048 from 2^20.  Every setting gives
 * the same trajectory.  Forces a list rebuild. */
int rh_engine3_tune(rh_engine3* e, int variant, uint32_t block);
/* The engine's HIP stream (hipStream_t), for callers that time or order work around it. */
int rh_engine3_stream(rh_engine3* e, void** stream);
/* Mean duration (ms) of the force+integrate launches of the last step() call, measured with HIP
 * events on the engine's stream (0 if not measured); enable with rh_engine3_set_timing(e, 1). */
int rh_engine3_set_timing(rh_engine3* e, int on);
int rh_engine3_kernel_ms(rh_engine3* e, double* force_ms, uint64_t* force_launches, double* rebuild_ms, uint64_t* rebuilds);

/* ---- the 2D engine: md::GpuEngine (periodic / hard walls, pair forces in f32) ---- */
typedef struct rh_engine2 rh_engine2;
int rh_engine2_create(const rh_system2* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, const rh_mode* mode, rh_engine2** out);
int rh_engine2_destroy(rh_engine2* e);
int rh_engine2_bytes_needed(const rh_system2* sys, uint64_t n, const rh_mode* mode, uint64_t* bytes);
int rh_engine2_device_bytes(const rh_engine2* e, uint64_t* bytes);
int rh_engine2_step(rh_engine2* e, uint64_t steps);
int rh_engine2_reverse(rh_engine2* e);
int rh_engine2_upload(rh_engine2* e, const int64_t* pos, const int32_t* vel, const uint32_t* species);
int rh_engine2_read_state(rh_engine2* e, int64_t* pos, int32_t* vel, uint32_t* species);
int rh_engine2_kicks(rh_engine2* e, int32_t* kicks, double* potential);
int rh_engine2_compute_forces(rh_engine2* e);                                                /* engine.rs:875-878 */
int rh_engine2_read_kinetic_total(rh_engine2* e, double* ke);
int rh_engine2_list_stats(rh_engine2* e, rh_list_stats* out);
/* md::Arith (crates/md/src/engine.rs:92-100; GpuEngine::with_options :609-625): 0 = f32 pair
 * forces, 1 = all-integer pair forces from md_core::table::IntTables (table.rs:17-105), which are
 * bit-identical on every device (pair forces only; either force mode). */
int rh_engine2_set_arith(rh_engine2* e, int arith);
int rh_engine2_tune(rh_engine2* e, int variant, uint32_t block);
int rh_engine2_stream(rh_engine2* e, void** stream);
int rh_engine2_n(const rh_engine2* e, uint64_t* n);                                          /* engine.rs:777-779 */
int rh_engine2_steps(const rh_engine2* e, int64_t* steps);                                   /* GpuEngine::steps, :595 */
int rh_engine2_update_system(rh_engine2* e, const rh_system2* sys);                           /* engine.rs:895-934 */
int rh_engine2_set_sort_interval(rh_engine2* e, uint32_t interval);                           /* sort_interval */
int rh_engine2_sort_now(rh_engine2* e);                                                       /* engine.rs:943-948 */
int rh_engine2_read_potential(rh_engine2* e, double* pe);                                     /* kicks().1, :1090-1097 */
int rh_engine2_read_positions(rh_engine2* e, int64_t* pos);                                   /* :1085-1088 */
int rh_engine2_kick_region(rh_engine2* e, const double lo[2], const double hi[2], double scale, const double add[2]); /* :1026-1038 */
int rh_engine2_set_species(rh_engine2* e, const uint32_t* ids, uint64_t count, uint32_t species);                /* :1042-1053 */
int rh_engine2_set_species_velocity(rh_engine2* e, uint32_t species, const double v[2]);                         /* :1056-1064 */
/* The ledger totals (:953-955) and per wall and compartment along it (Buffers::
 * read_compartment_ledger, :455-470): energy / impulse / hits are [7][per_wall] arrays. */
int rh_engine2_read_ledger(rh_engine2* e, rh_ledger2* out);
int rh_engine2_read_compartment_ledger(rh_engine2* e, uint32_t per_wall, double* energy, double* impulse, uint64_t* hits);
/* The piston (:958-961): out = {height sigma, velocity sigma/tau, gas force on it eps/sigma,
 * work done on it by the gas eps}; set_piston_speed (:937-940) for a driven piston (sigma/tau). */
int rh_engine2_read_piston(rh_engine2* e, double out[4]);
int rh_engine2_set_piston_speed(rh_engine2* e, double u);
/* XOR one bit of a coordinate: of the particle with id `particle` (:964-973), or of every
 * particle whose species is in the bitmask `species` (:1014-1023).  bit < 64; a flip that would
 * put a particle outside the box (every bit from 52 up, for any box the engine accepts) is
 * refused with RH_ERR_INVALID and changes nothing. */
int rh_engine2_flip_bit(rh_engine2* e, uint32_t particle, uint32_t axis, uint32_t bit);
int rh_engine2_flip_species_bit(rh_engine2* e, uint32_t species, uint32_t axis, uint32_t bit);
/* Bonds at run time (:976-1011): *changed = 1 if the bond was added / removed. */
int rh_engine2_add_bond(rh_engine2* e, uint32_t a, uint32_t b, uint32_t type, int* changed);
int rh_engine2_remove_bond(rh_engine2* e, uint32_t a, uint32_t b, int* changed);

/* ---- the batched engine (no reference counterpart: the reference runs one system per engine;
 * its 2D "compartments", crates/md-core/src/system.rs:41-48, are the nearest prior art) ----
 * Independent 2D and 3D systems of any sizes share one device arena and advance together, one
 * launch per kernel for all of them, in hipGraph-captured chunks.  Systems may be added and
 * removed between runs (continuous batching).  Each system's trajectory is bit-identical to
 * running it alone in rh_engine3 / rh_engine2. */
typedef struct rh_batch rh_batch;
/* budget_bytes: device memory for the arenas (0: half the free memory); max_systems: slots. */
int rh_batch_create(uint64_t budget_bytes, uint32_t max_systems, rh_batch** out);
int rh_batch_destroy(rh_batch* b);
/* Admits a system with `steps` steps to run, Verlet skin `skin`; *id = -1 when it does not fit
 * in the free arena space now (the admission rule).  Systems may have hard, soft and heat-bath
 * walls, 3D ones the Andersen bath and 2D ones the thermostat region (fan, bath layer), all with
 * the ledger booked as the single-system engines book it; compartments, the membrane, the piston
 * and bonds are refused. */
int rh_batch_add3(rh_batch* b, const rh_system3* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, double skin, uint64_t steps, int32_t* id);
int rh_batch_add2(rh_batch* b, const rh_system2* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, double skin, uint64_t steps, int32_t* id);
int rh_batch_remove(rh_batch* b, int32_t id);
int rh_batch_set_steps(rh_batch* b, int32_t id, uint64_t steps);
/* Enqueues `iterations` iterations (every unfinished system takes one step per iteration). */
int rh_batch_run(rh_batch* b, uint64_t iterations);
int rh_batch_sync(rh_batch* b);
/* Synchronises; remaining[id] = steps left for each id < count (0 for unused ids). */
int rh_batch_poll(rh_batch* b, uint64_t* remaining, uint32_t count);
int rh_batch_read_state(rh_batch* b, int32_t id, int64_t* pos, int32_t* vel, uint32_t* species);
int rh_batch_list_stats(rh_batch* b, int32_t id, rh_list_stats* out);
/* The system's ledger since it joined the batch (GpuEngine3::read_ledger, engine3.rs:645-647, and
 * GpuEngine's for 2D, whose entries 0-3 are the walls): energy and impulse per wall (2 axis +
 * side) and for the bath (entry 6), and hits.  The same totals as the system run alone. */
int rh_batch_read_ledger(rh_batch* b, int32_t id, rh_ledger3* out);
/* Iterations per captured hipGraph (0: launch directly). */
int rh_batch_set_graph_chunk(rh_batch* b, uint32_t chunk);
/* Tile settings for systems added from now on, per dimension and size class (floor(log2 n)):
 * list variant (1 half list within work-groups, the default, built in 3D over a row-ordered
 * copy where the box allows; 2 the same half list built over the cell stencil; 0 every pair at
 * both ends), cell occupancy (cells at least cell_factor x (rc + skin) wide,
 * 1 <= cell_factor <= 4) and block
 * (slots per half-list block and step work-group: 256, the default, 512 or 1024; the system's
 * slots are whole blocks).  Results are identical for every setting; micrographia_amd.autotune
 * picks them by timing. */
int rh_batch_set_tuning(rh_batch* b, uint32_t dim, uint32_t size_class, uint32_t variant, double cell_factor, uint32_t block);
/* The batch's stream: every iteration's work (its side branch included) is joined back into it,
 * so work queued on it after rh_batch_run follows the whole run. */
int rh_batch_stream(rh_batch* b, void** stream);
int rh_batch_capacity(rh_batch* b, uint64_t* slots, uint64_t* list_entries, uint64_t* cells, uint32_t* active);
int rh_batch_bytes_needed3(rh_batch* b, const rh_system3* sys, uint64_t n, double skin, uint64_t* bytes);
int rh_batch_bytes_needed2(rh_batch* b, const rh_system2* sys, uint64_t n, double skin, uint64_t* bytes);

/* ---- gpu-prims on device buffers (crates/gpu-prims/src/prims.rs:116-312, threefry.rs) ----
 * Pointers are device memory; stream may be NULL (the null stream).  These run synchronously
 * with respect to the host (they synchronise the stream before returning). */
int rh_prims_exclusive_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream);
int rh_prims_radix_sort_u32(uint32_t* d_keys, uint32_t* d_vals, uint64_t n, uint32_t key_bits, void* stream);
/* No counterpart in the reference: a variant of rh_prims_radix_sort_u32 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc/hipFree: the caller's arrays are one buffer of the
 * sort's ping-pong pair, and a fixed device array inside the library (about 8.4 MB, held by each
 * GPU it runs on until the process ends) holds the other buffer, the digit histograms and the
 * scan's block sums.  n <= RH_SORT_NOALLOC_MAX, else RH_ERR_INVALID and nothing runs.  Same
 * result as rh_prims_radix_sort_u32 (stable, sorted in place), but enqueued on `stream` without
 * synchronising: the caller synchronises before reading.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (other streams, other threads) race. */
#define RH_SORT_NOALLOC_MAX (1u << 20)
int rh_prims_radix_sort_u32_noalloc(uint32_t* d_keys, uint32_t* d_vals, uint64_t n, uint32_t key_bits, void* stream);
int rh_prims_reduce_sum_f64(const double* d_in, double* d_out, uint64_t n, void* stream);
int rh_prims_threefry(uint32_t* d_out, uint64_t n, uint32_t ctr0, uint32_t ctr1, uint32_t key0, uint32_t key1, void* stream);
int rh_prims_atomic_add_i64(const int64_t* d_vals, uint64_t n, int64_t* d_acc, void* stream);
/* The engine's float -> integer kick conversion on device: out[i] = i32(rint(x[i])), saturating,
 * NaN -> 0 (Rust `as i32` after round_ties_even). */
int rh_prims_kick_convert(const float* d_x, int32_t* d_out, uint64_t n, void* stream);
/* The stochastic boundaries' draws on the device (md_core::three::andersen and
 * thermal_velocity, three.rs:232-265, as the engines compute them with the device's f64
 * log/cos/sin), for keys (id0 + k, clock), k < count: d_out[6k..6k+6) = the Andersen bath's
 * three velocities and a low heat-bath wall's normal and two tangential ones at temperature t
 * for mass m, each as the integer displacement per step (round(v dt 2^32), saturating), every
 * key redrawn.  A check of those draws against the host C library's. */
int rh_prims_stochastic_draws(uint64_t seed, double t, double mass, double dt, uint32_t id0, uint64_t count, uint32_t clock, int32_t* d_out, void* stream);
/* The pair kick divides by r2 twice with one shared, refined reciprocal when r2 >= r2_fast
 * (derived per species pair); below it, IEEE division.  rh_pair_division_floor gives r2_fast for
 * a pair; rh_prims_check_pair_division compares the shared-reciprocal result with IEEE division
 * for EVERY f32 value of r2 in [r2_fast, rc^2) on the device and counts differing results. */
int rh_pair_division_floor(double eps, double sigma, double rc, float* r2_fast);
int rh_prims_check_pair_division(double eps, double sigma, double rc, uint64_t* checked, uint64_t* mismatches, float* first_bad, void* stream);

#ifdef __cplusplus
}
#endif
#endif
