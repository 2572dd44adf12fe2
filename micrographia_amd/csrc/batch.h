// Batched engine: many independent systems (2D and 3D, any sizes) share one device arena and
// advance together, one launch per kernel for all of them.  Systems join and leave between runs
// (continuous batching): their slot, list and cell ranges come from first-fit allocators over
// arenas sized by a device-memory budget, which is also the admission rule.
#pragma once

#include <map>
#include <vector>

#include "engine.h"
#include "range_alloc.h"

namespace rh {

constexpr uint32_t BATCH_WG = 256;         // slots per work-group of the per-slot kernels, the slot unit
constexpr uint32_t BATCH_MAX_BLOCK = 1024;  // largest half-list block (slots per step work-group)
// Build kinds: 0-1 2D, 2-3 3D over the cell stencil, 4-5 3D over the row-ordered copy (+1: more
// than one species); the rebuild lists hold one list per kind, then one of all kinds together.
constexpr uint32_t BUILD_KINDS = 6;
constexpr uint32_t RB_ALL = BUILD_KINDS;

// Per-system device descriptor.
struct BSys {
    uint32_t active, dim, n, slot0, nslots;
    uint32_t nc[3];
    // Cell keys: tiles of 2^tile_sh cells a side in linear order (nt[] tiles per axis, the last
    // ones padded), Morton order inside a tile, so a block of consecutive slots covers a compact
    // region and keeps more of its pairs inside the block.  ncells counts the padded keys.
    uint32_t tile_sh, nt[3];
    uint32_t ncells, cell0;
    // The row-ordered copy for the 3D list build (rows != 0): each (y, z) cell's row of rows_nxf
    // fine x bins, with rows_g ghost bins at both ends on a periodic x axis (rows_ne = nxf + 2 g
    // per row).  The fine bins' counts and the extended bins' starts follow the cells in the
    // system's cell range (cell_count and cell_start from cell0 + ncells); the copy itself (every
    // slot once, plus at most one ghost) is at items + 2 slot0.  cell_words: the whole cell range.
    uint32_t rows, rows_nxf, rows_g, rows_ne, cell_words;
    float rows_inv_wx, rows_reach;
    double rows_xscale;
    uint32_t periodic, nspecies, pinned, soft_walls, range_check, cap;
    uint32_t variant;  // 1: half list within work-groups (default; built over the rows in 3D), 2: the same built over the cell stencil, 0: every pair at both ends
    uint32_t block;    // slots per half-list block and step work-group: 256, 512 or 1024
    uint32_t bkind;    // build kind: 0 (2D), 2 (3D, cell stencil) or 4 (3D, rows), + (species > 1)
    uint64_t nbr0;  // list entry k of local slot i at nbr0 + k * nslots + i
    int64_t box[3], half[3];
    double boxf[3], inv_box[3];  // positions are exact f64 integers (2^-32 sigma units)
    double box16[3], inv_box16[3];  // box / 16 and 16 / box: the image on positions in sixteenths
    double face16;                  // a block this far (sixteenths) from every periodic face needs no image
    double cell_scale[3], cell_w64[3];
    float cell_w[3];
    float kick_c[4];
    int32_t gravity[3];
    float half_skin2;
    // The ledger (walls and bath) is booked when ledger != 0: some axis walled, or the bath.
    // Stochastic boundaries (3D only): the Andersen bath and heat-bath walls.
    uint32_t ledger, bath, bath_species, thermal_walls, seed_lo, seed_hi;
    double dt, bath_p;
    double mass[4];        // per species, for the energies the ledger books
    // the 2D thermostat region: along bath_axis in [bath_lo, bath_hi), the fan's drift along x
    uint32_t bath_axis;
    int32_t bath_drift;
    double bath_lo, bath_hi;
    double bath_sd[4];     // sqrt(T / m) of the bath per species (sigma/tau)
    double wall_sd[6][4];  // sqrt(T / m) per heat-bath wall and species
    PairTable pt;
};

// Per-system control words (device).
struct BCtl {
    uint32_t stale, stale_next, overflow, remaining, ppos, pmeta, do_rebuild, maxcount;
    uint32_t clock, rebuilds, done_lo, done_hi, skipped, pad0, pad1, pad2;
};

struct BBufs {
    double* pos[2][3];  // exact integers below 2^53 (2^-32 sigma units)
    int32_t* vel[2][3];
    uint32_t* species[2];
    uint32_t* ids[2];
    double* ref[3];
    double4* aos[2];  // positions in sixteenths packed per slot (x, y, z or 0, 0), for the step's neighbour gathers
    uint32_t* wave_chunks;  // in-block and cross-block list chunks in each 64-slot wave's region (2 per wave)
    float4* rel;  // position relative to its cell's corner (sigma, f32), for list builds
    float4* items;  // the row-ordered copies (x absolute, y and z cell-relative, local slot)
    uint32_t* key;
    uint32_t* rank;
    uint32_t* nbr;
    uint32_t* cell_count;
    uint32_t* cell_start;
    uint32_t* cell_end;
    const uint32_t* wg_sys;  // system of each BATCH_WG-slot unit (~0u: none)
    // Work lists.  step_list: per kind of step launch, the block indices of its systems (written
    // by the host with the tables).  rb_list: per build kind, and for all kinds together, the
    // units of the systems rebuilding this iteration, appended by k_bb_decide; rb_count holds
    // their lengths (build kinds, then all) and is cleared at the end of each iteration.
    const uint32_t* step_list;
    uint32_t* rb_list;
    uint32_t* rb_count;
    const BSys* sys;
    BCtl* ctl;
    unsigned long long* ledger;  // LEDGER_WORDS per system slot (fixed point, 2^-32)
};

struct BatchMember {
    SystemDesc sys;
    uint64_t n;
    uint32_t slot0, nslots, cell0, ncells, cap;
    uint64_t nbr0, nbr_entries;
    double skin;
    bool overflowed = false;  // its list ever overflowed (and was regrown, or is waiting to be)
};

class Batch {
   public:
    // budget_bytes: device memory the arenas may use (0: a share of the free memory).
    Batch(uint64_t budget_bytes, uint32_t max_systems);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // Admits a system if its ranges fit now; returns its id, or -1 (not admitted).
    int add(const SystemDesc& sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, double skin, uint64_t steps);
    void remove(int id);
    void set_steps(int id, uint64_t steps);
    void run(uint64_t iterations);
    void sync();
    // Copies every system's control words to the host (synchronises); returns remaining steps.
    void poll(uint64_t* remaining, uint32_t count);
    void read_state(int id, int64_t* pos, int32_t* vel, uint32_t* species);
    // The system's ledger (walls and bath) since it joined: 7 entries of energy, impulse, hits.
    void read_ledger(int id, double energy[7], double impulse[7], uint64_t hits[7]);
    rh_list_stats stats(int id);
    void set_graph_chunk(uint32_t chunk) { graph_chunk_ = chunk; }
    hipStream_t stream() const { return stream_; }
    uint64_t bytes_needed(int dim, uint64_t n, const SystemDesc& sys, double skin) const;
    uint64_t free_slots() const { return slots_.capacity() - slots_.used(); }
    uint32_t active() const { return (uint32_t)members_.size(); }
    void budget(uint64_t* slot_cap, uint64_t* list_cap, uint64_t* cell_cap) const;
    // Tile settings for systems added from now on, per dimension and size class (floor(log2 n)):
    // the list variant, the cell occupancy (cells at least cell_factor x (rc + skin) wide) and
    // the half-list block (slots per step work-group; a system's slots start at a multiple of
    // it and fill whole blocks).
    void set_tuning(int dim, uint32_t size_class, uint32_t variant, double cell_factor, uint32_t block);

   private:
    struct Tuning {
        uint32_t variant = 1;
        double cell_factor = 1.0;
        uint32_t block = BATCH_WG;
    };
    Tuning tuning_[2][32];
    void plan(const SystemDesc& sys, uint64_t n, double skin, BSys& d, uint32_t* cap) const;
    void upload_tables();
    void enqueue_iteration();
    template <uint32_t BS>
    void launch_steps(uint32_t b, hipStream_t st, int dim);
    void handle_overflow();
    void invalidate_graph();

    hipStream_t stream_ = nullptr;
    hipStream_t aux_ = nullptr;  // an iteration's side branch (small 3D systems' builds, 2D systems)
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr, ev_join2_ = nullptr;  // the branches of an iteration
    BBufs B_;
    BSys* d_sys_ = nullptr;
    BCtl* d_ctl_ = nullptr;
    uint32_t* d_wg_ = nullptr;
    uint32_t max_sys_;
    RangeAlloc slots_, lists_, cells_;  // units: BATCH_WG slots, LIST_UNIT entries, CELL_UNIT cells
    std::map<int, BatchMember> members_;
    std::vector<BSys> h_sys_;
    std::vector<int> free_ids_;
    uint32_t grid_wgs_ = 0;  // work-groups covering the slot high-water mark
    uint32_t* d_step_list_ = nullptr;
    uint32_t* d_rb_list_ = nullptr;
    uint32_t* d_rb_count_ = nullptr;
    uint32_t units_cap_ = 0;        // BATCH_WG-slot units in the slot arena
    uint32_t step_off_[24] = {}, step_cnt_[24] = {};  // step launch kind 8 b + kind: its list
    uint32_t rb_off_[RB_ALL + 1] = {}, rb_cap_[RB_ALL + 1] = {};  // rebuild lists: offset, room (units)
    bool tables_dirty_ = true;
    uint32_t graph_chunk_ = 100;
    hipGraphExec_t graph_ = nullptr;
    uint32_t graph_wgs_ = 0, graph_len_ = 0;
    BCtl* h_ctl_ = nullptr;
    std::vector<void*> allocs_;
    void* dalloc(size_t bytes);
};

}  // namespace rh
