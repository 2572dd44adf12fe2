// The Rahman engine on one GPU: integer positions (i64 per axis, 2^-32 sigma units), integer
// velocities (i32 displacement per step), f32 Lennard-Jones pair forces rounded per pair to
// integer kicks, summed with wrapping i32 adds (equal to the low 32 bits of an i64 sum, so the
// total does not depend on summation order).
#pragma once

#include <vector>

#include "core.h"
#include "prims.h"

namespace rh {

// Control words shared by the device-gated launches (all u32).
enum : uint32_t {
    CTL_STALE = 0,       // the Verlet list must be rebuilt before the next step
    CTL_STALE_NEXT = 1,  // set by the running step when some particle moved half the skin
    CTL_OVERFLOW = 2,    // a list build overflowed its capacity (steps stop until regrown)
    CTL_CLOCK = 3,       // steps taken in either direction (stochastic boundaries)
    CTL_REBUILDS = 4,
    CTL_REMAINING = 5,   // steps still to take in the current step() call
    CTL_PPOS = 6,        // which position buffer is current (they alternate every step)
    CTL_PMETA = 7,       // which velocity/species/id buffer is current (alternate per rebuild)
    CTL_DO_REBUILD = 8,  // the gate of the rebuild chain launched at the head of a chunk
    CTL_MAXCOUNT = 9,    // largest list at the last build
    CTL_DONE_LO = 10,
    CTL_DONE_HI = 11,
    CTL_SKIPPED = 12,    // gated step launches that found the list stale
    CTL_FLIP_BAD = 13,   // bit flips that would have left the box (flip_bit's check pass)
    CTL_WORDS = 32
};

// f32 pair constants per species pair [a*4 + b] and kick factors per species, kept in device
// memory so per-lane species indexing needs no private copy.
struct PairTable {
    float rc2[16], eps4[16], six_eps4[16], sigma2[16], shift[16], list_r2[16];
    float kick_c[4];
    float r2_fast[16];  // pairs with r2 >= this use the shared-reciprocal division (0: never)
    // f64 constants for the energy readout of the all-integer forces (Pair::energy in f64)
    double eps4_64[16], sigma2_64[16], rc2_64[16], shift_64[16];
};

template <int D>
struct Params {
    uint32_t n;
    uint32_t nc[3];
    uint32_t mbits;      // bits per axis of the Morton cell key
    uint32_t periodic;   // bit a: axis a periodic
    uint32_t nspecies;
    uint32_t pinned;
    uint32_t soft_walls; // bit 2a+side: a soft wall
    uint32_t cap;
    int64_t box[3];
    int64_t half[3];
    double boxf[3];        // box length as f64 (positions are kept as exact f64 integers)
    double inv_box[3];     // 1 / box on periodic axes, 0 on walled ones
    double box16[3];       // box / 16 and 16 / box: the same image on positions held in sixteenths
    double inv_box16[3];
    double face16;         // a block this far (sixteenths) from every periodic face needs no image
    double cell_scale[3];  // cells per position unit
    float kick_c[4];
    int32_t gravity[3];
    float half_skin2;
    uint32_t block;        // half-list variant: particles per work-group block
    // Row-ordered list build (3D half list): rows are the (y, z) cells, each binned into
    // rows_nxf fine cells along x, with rows_g ghost bins at each end on a periodic x axis
    // (rows_ne bins a row in all); rows != 0 when the geometry allows it.
    uint32_t rows;
    uint32_t rows_nxf;
    uint32_t rows_g;
    uint32_t rows_ne;
    double rows_xscale;    // fine x cells per position unit
    float rows_inv_wx;     // fine x cells per sigma
    float rows_reach;      // list radius of the widest pair plus a rounding margin (sigma)
    uint32_t range_check;  // list separations may reach 8 sigma: keep the reference's range test
    float cell_w[3];       // cell width (sigma), f32
    double cell_w64[3];    // cell width (sigma)
    // Stochastic boundaries and the ledger (int_engine3.rs:138-206, three.rs:232-265)
    uint32_t bath;           // Andersen heat bath on
    uint32_t bath_species;   // bit s: species s is coupled to the bath
    uint32_t thermal_walls;  // bit 2a+side: a heat-bath wall
    uint32_t ledger;         // book wall and bath exchanges (some axis walled, or a bath)
    uint32_t seed_lo, seed_hi;
    double bath_p;           // redraw probability per particle per step (rate * dt)
    double bath_sd[4];       // sqrt(T / m) of the bath per species (sigma/tau)
    double wall_sd[6][4];    // sqrt(T / m) per heat-bath wall and species
    double dt;
    double mass[4];
    // All-integer pair forces from tables (md::Arith::Int, 2D only)
    uint32_t arith;          // 0: f32 pair forces, 1: integer tables
    uint32_t itab_len;       // entries per species pair
    // md_core::System's 2D boundaries (system.rs:26-61).  feat selects the boundary path: pair
    // kicks, then the per-particle terms (bonds, soft walls, piston), then the leapfrog with
    // every boundary, each its own launch.
    uint32_t feat;
    uint32_t comp_filter;    // insulating compartments: pairs in different ones do not interact
    uint32_t comp_shift[2];  // log2 of the compartment side (position units); 0: none
    double comp_scale[2];    // 2^-shift, or 0 on an axis without compartments
    uint32_t bath_axis;      // the thermostat acts where bath_lo <= x[bath_axis] < bath_hi
    int32_t bath_drift;      // the fan's drift as a displacement per step (x)
    double bath_lo, bath_hi;
    uint32_t membrane, mem_axis, mem_blocks, pore_shift;
    double mem_pos;          // the membrane's position (position units)
    int64_t pore_mask, pore_width;
    uint32_t piston, piston_free;
    double piston_mass, piston_load;  // Free drive: mass and the external force (pushing down)
    int64_t piston_lo, piston_hi;     // its stops (position units)
    uint32_t bonds;          // a partner table is present
    uint32_t bond_fene;      // bit t: bond type t is FENE (else harmonic)
    float bond_k[4], bond_r0[4];
    PairTable pt;
};

// The piston's state on the device (GpuEngine's ctl words PISTON_*, engine.rs:57-66).
struct PistonDev {
    long long w;                   // height (position units)
    int32_t v;                     // displacement per step
    int32_t pad;
    unsigned long long force_sum;  // the gas force of the current evaluation, 2^-16 eps/sigma (i64)
    double force;                  // the gas force of the last step (eps/sigma)
    double work;                   // work done on it by the gas (eps)
};

template <int D>
struct Bufs {
    // Positions in 2^-32 sigma units, held as f64: integers below 2^53, so exact (the C ABI
    // takes and returns i64).
    double* pos[2][D];   // [parity][axis]
    // The same positions packed per particle (x, y, z, 0 / x, y), written beside the SoA copy by
    // the reorder and the half-list step: a neighbour gather reads one 32-byte record (one cache
    // line) instead of one 8-byte word from each axis array.
    double* aos[2];
    // Half list: a stream of 16-byte chunks in B.nbr, each 64-particle wave of the build filling
    // a region of 64 cap/4 slots, in-block chunks (up to eight offsets, pack_inblock) from its
    // start and cross-block chunks (up to four slots, tag_chunk) from its end down; per wave the
    // number of each (wave_chunks[2 w], wave_chunks[2 w + 1]).
    uint32_t* wave_chunks;
    int32_t* vel[2][D];
    uint32_t* species[2];
    uint32_t* ids[2];
    double* ref[D];      // positions at the last list build
    float4* rel;         // position relative to its cell's corner (sigma, f32; w unused), for list builds
    // Row-ordered copy for the row build: per entry (x sigma, y and z relative to the cell's
    // corner, slot | species << 30 with several species, else the slot), ordered by row and
    // fine x bin.  row_cnt[r * nxf + b] counts the slots in bin b of row r and row_rank is a
    // slot's place among them; row_start[r * ne + e] is the first entry of extended bin e of row
    // r (a dense exclusive scan), whose bins are the row's ghost copies of its last rows_g bins
    // (x - box), its own bins, then ghost copies of its first rows_g bins (x + box).
    float4* items;
    uint32_t* row_cnt;
    uint32_t* row_start;
    uint32_t* row_rank;
    uint32_t* nbr;       // column-major: entry k of slot i at k*n + i
    uint32_t* nbr_count;
    uint32_t* cell_start;
    uint32_t* cell_end;
    int32_t* kicks[D];   // kicks of the last kicks() / cell-list step (slot order)
    double* pe;          // per-slot potential energy (halved pair sums + walls)
    double* ke;          // per-slot kinetic energy
    uint32_t* ctl;
    const PairTable* pt;
    unsigned long long* ledger;  // [7][3]: energy, impulse (2^-32 units), hits per wall; 6 = bath
    const int32_t* itab;         // integer force tables [(a * ns + b) * len + k] (arith 1)
    const uint32_t* irc2;        // their cutoffs [a * ns + b], 2^-24 sigma^2
    // 2D boundary path
    const uint32_t* partners;    // per particle id: 4 x (partner id, bond type), UINT32_MAX padded
    uint32_t* slot_of;           // particle id -> slot (refreshed at every reorder)
    unsigned long long* cledger; // [7][256][3]: the ledger per wall and compartment along it
    PistonDev* piston;
};

// The smallest r2 of a species pair from which the shared-reciprocal division of the pair kick
// provably equals IEEE division (see pair_fr); +inf if none.
float fast_division_floor(float rc2, float sigma2, float six_eps4);

struct SystemDesc {
    int dim;
    int64_t box_len[3];
    double dt;
    uint32_t nspecies, pinned;
    double masses[4];
    rh_pair pairs[16];
    uint32_t walled[3];
    uint32_t wall_kind[3][2];
    double wall_t[3][2];
    double gravity[3];
    uint64_t seed;
    uint32_t has_bath, bath_species;
    double bath_t, bath_rate;
    // 2D only (md_core::System, system.rs:26-61); see rh_system2
    uint32_t compartments[2];
    uint32_t diathermal;
    uint32_t bath_axis;
    double bath_drift;
    int64_t bath_region[2];
    uint32_t has_membrane, membrane_axis, membrane_blocks, pore_shift;
    int64_t membrane_position, pore_width;
    uint32_t has_piston, piston_free;
    double piston_speed, piston_mass, piston_force, piston_start, piston_min, piston_max;
    uint32_t nbond_types;
    uint32_t bond_fene[4];
    double bond_k[4], bond_r0[4];
    std::vector<uint32_t> bonds;  // [id_a, id_b, type] per bond
};

template <int D>
class Engine {
   public:
    Engine(const SystemDesc& sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, const rh_mode& mode);
    // A plan only: the constructor's checks and parameters, no device work (planned_bytes()).
    Engine(const SystemDesc& sys, uint64_t n, const rh_mode& mode);
    uint64_t planned_bytes() const;
    uint64_t alloc_bytes() const { return alloc_bytes_ + sort_bytes_; }
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    uint64_t n() const { return n_; }
    int64_t steps() const { return steps_; }
    void step(uint64_t steps);
    void reverse();
    void update_system(const SystemDesc& sys);
    void upload(const int64_t* pos, const int32_t* vel, const uint32_t* species);
    void read_state(int64_t* pos, int32_t* vel, uint32_t* species);
    void kicks(int32_t* kicks, double* potential);
    double kinetic_total();
    void measure_kinetic();
    void read_kinetic(float* ke, uint32_t* species);
    void compute_forces();
    double potential();
    void kick_region(const double lo[3], const double hi[3], double scale, const double add[3]);
    void set_species(const uint32_t* ids, uint64_t count, uint32_t species);
    void set_species_velocity(uint32_t species, const double v[3]);
    rh_list_stats list_stats();
    void read_ledger(rh_ledger3* out);
    // 2D boundaries (md::GpuEngine, engine.rs:937-1023)
    void read_compartment_ledger(uint32_t per_wall, double* energy, double* impulse, uint64_t* hits);
    void read_piston(double out[4]);
    void set_piston_speed(double u);
    void flip_bit(uint32_t particle, uint32_t axis, uint32_t bit);
    void flip_species_bit(uint32_t species, uint32_t axis, uint32_t bit);
    bool add_bond(uint32_t a, uint32_t b, uint32_t type);
    bool remove_bond(uint32_t a, uint32_t b);
    void sort_now();
    void read_positions(int64_t* pos);
    // Pair-force arithmetic (md::Arith, engine.rs:92-100): 0 f32 pair forces, 1 all-integer
    // tables (2D, cell-list mode).
    void set_arith(int arith);
    hipStream_t stream() const { return stream_; }
    void set_timing(bool on) { timing_ = on; }
    void kernel_ms(double* force_ms, uint64_t* force_launches, double* rebuild_ms, uint64_t* rebuilds) const;
    void set_sort_interval(uint32_t v) { sort_interval_ = v; }
    // Kernel variant: 0 full Verlet list (one lane per particle, both ends of every pair), 1 half
    // list within work-group blocks of `block` particles (pairs in a block once, the partner's kick
    // accumulated in LDS) plus the full list across blocks, built from the row-ordered copy where
    // the geometry allows it; 2 the same half list built over the cell stencil.
    void tune(int variant, uint32_t block);
    int variant() const { return variant_; }
    uint32_t block() const { return P_.block; }

   private:
    void build_params(const SystemDesc& sys, const rh_mode& mode);
    void alloc_list(uint32_t cap);
    void enqueue_rebuild(bool gated);
    void enqueue_cells_at_current(const uint32_t* gate);  // keys + sort + ranges of current slots
    void enqueue_force_cells(const uint32_t* gate);       // kicks/pe of current slots
    void step_list(uint64_t steps);
    void step_cells(uint64_t steps);
    void read_ctl(uint32_t* out);
    void write_ctl(uint32_t word, uint32_t value);
    void handle_overflow();
    void upload_tables();
    void record_timer(bool rebuild, hipEvent_t start);
    hipEvent_t timer_event();
    void collect_timers();
    void launch_step();
    void launch_boundary_step(uint32_t clock_add, bool gated);  // the 2D boundary path's step
    // The split path (pair kicks, per-particle terms, leapfrog as separate launches over a full
    // list): 2D systems with boundaries, and integer forces in Verlet mode.
    bool split_path() const { return P_.feat != 0u || P_.arith != 0u; }
    // The half-list kernels run unless the split path does, or list separations may reach the
    // 8 sigma range limit (then the full-list kernel keeps the reference's range test).
    // (Half-list chunks carry their tag in the top four bits of each index: below 2^28 particles.)
    bool half_path() const { return variant_ >= 1 && !split_path() && P_.range_check == 0u && n_ < (uint64_t(1) << 28); }
    // The half list built from the row-ordered copy (variant 1 where the geometry allows it;
    // variant 2 keeps the cell-stencil build).
    bool row_path() const { return D == 3 && variant_ == 1 && half_path() && P_.rows != 0u; }
    void plan_rows();
    void alloc_rows();
    uint64_t row_keys_ = 0;   // extended bins of all rows (row_start has one more entry)
    uint64_t row_bins_ = 0;   // main bins of all rows (row_cnt)
    void* row_scan_temp_ = nullptr;
    void* key_scan_temp_ = nullptr;  // the row path's counting sort of the cell keys
    void enqueue_extras(uint32_t flags);
    void write_partners();
    void flip(uint32_t sel, int by_species, uint32_t axis, uint32_t bit);

    SystemDesc sys_;
    rh_mode mode_;
    Params<D> P_;
    Bufs<D> B_;
    uint64_t n_ = 0;
    int64_t steps_ = 0;
    uint32_t key_bits_ = 0;
    uint64_t ncell_keys_ = 0;
    RadixSort sort_;
    hipStream_t stream_ = nullptr;
    uint32_t* h_ctl_ = nullptr;  // pinned readback slots
    std::vector<hipEvent_t> ev_;
    void* reduce_temp_ = nullptr;
    double* d_scalar_ = nullptr;
    uint32_t* d_set_list_ = nullptr;
    uint64_t set_list_cap_ = 0;
    double lifetime_ = 8.0;  // running estimate of steps per list
    bool timing_ = false;
    // Pairs of events around timed launches, resolved after the stream drains.
    std::vector<hipEvent_t> ev_pool_;
    size_t ev_used_ = 0;
    struct Timed {
        hipEvent_t a, b;
        bool rebuild;
    };
    std::vector<Timed> timed_;
    uint32_t since_sort_ = 100;
    PairTable* d_pt_ = nullptr;
    double force_ms_ = 0.0, rebuild_ms_ = 0.0;
    uint64_t force_launches_ = 0, rebuild_launches_ = 0;
    uint32_t sort_interval_ = 100;
    int variant_ = 1;
    uint32_t block_ = 512;
    std::vector<void*> allocs_;
    uint64_t alloc_bytes_ = 0, sort_bytes_ = 0, list_bytes_ = 0;
    void* dalloc(size_t bytes);
    std::vector<uint32_t> partners_h_;  // the bond partner table (bonds may change at run time)
    void alloc_boundary_buffers();
    void upload_int_tables();
};

}  // namespace rh
