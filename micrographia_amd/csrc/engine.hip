// Device kernels and the host driver of the Rahman engine (see engine.h).
//
// One step, Verlet-list path (crates/md/shaders/md3.wgsl force_list + integrate_list, restated):
//   k_step_list   one lane per particle, Morton-sorted slots: gather the listed neighbours'
//                 i64 positions, f32 pair math in md-ref's operation order, per-pair rint and a
//                 saturating convert to an i32 kick, wrapping i32 sum; then the leapfrog
//                 (v += kick + g; x += v; wrap / mirror) and the half-skin check, fused in one
//                 launch.  Positions alternate between two buffers (read one, write the other).
//   k_step_commit one lane: advance the control words (parity, remaining, stale flag).
// A chunk of K steps is preceded by a rebuild chain gated on the device's own stale flag:
//   keys (Morton cell) -> stable radix sort -> gather (reorder storage) -> cell ranges ->
//   k_build_list -> commit.  Steps after the list went stale are skipped on the device and
//   re-issued by the host, so every step runs with a valid list and the trajectory is the same
//   bit for bit whatever the chunking.
#include "engine.h"
#include "md_device.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <type_traits>

namespace rh {

constexpr int TPB = 256;

// ---------------------------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------------------------



// ---------------------------------------------------------------------------------------------
// Control kernels (one lane)
// ---------------------------------------------------------------------------------------------

__global__ void k_add_remaining(uint32_t* ctl, uint32_t add) { ctl[CTL_REMAINING] += add; }

__global__ void k_set_word(uint32_t* ctl, uint32_t word, uint32_t value) { ctl[word] = value; }

__global__ void k_rebuild_decide(uint32_t* ctl) {
    uint32_t go = (ctl[CTL_STALE] != 0u && ctl[CTL_REMAINING] != 0u && ctl[CTL_OVERFLOW] == 0u) ? 1u : 0u;
    ctl[CTL_DO_REBUILD] = go;
    if (go) ctl[CTL_MAXCOUNT] = 0u;
}

// Unconditional reorder (cell-list mode's periodic Morton sort).
__global__ void k_rebuild_force(uint32_t* ctl) {
    ctl[CTL_DO_REBUILD] = 1u;
    ctl[CTL_MAXCOUNT] = 0u;
}

__global__ void k_rebuild_commit(uint32_t* ctl) {
    if (ctl[CTL_DO_REBUILD]) {
        ctl[CTL_STALE] = 0u;
        ctl[CTL_STALE_NEXT] = 0u;
        ctl[CTL_PPOS] ^= 1u;
        ctl[CTL_PMETA] ^= 1u;
        ctl[CTL_REBUILDS] += 1u;
        ctl[CTL_DO_REBUILD] = 0u;
    }
}

__device__ __forceinline__ bool step_runs(const uint32_t* ctl) {
    return (ctl[CTL_STALE] | ctl[CTL_OVERFLOW]) == 0u && ctl[CTL_REMAINING] != 0u;
}

__global__ void k_step_commit(uint32_t* ctl) {
    if (step_runs(ctl)) {
        ctl[CTL_PPOS] ^= 1u;
        ctl[CTL_REMAINING] -= 1u;
        uint32_t lo = ctl[CTL_DONE_LO] + 1u;
        ctl[CTL_DONE_LO] = lo;
        if (lo == 0u) ctl[CTL_DONE_HI] += 1u;
        ctl[CTL_CLOCK] += 1u;
        ctl[CTL_STALE] = ctl[CTL_STALE_NEXT];
        ctl[CTL_STALE_NEXT] = 0u;
    } else if (ctl[CTL_REMAINING] != 0u) {
        ctl[CTL_SKIPPED] += 1u;
    }
}

// Cell-list steps are host-counted: advance clock and done by `steps`.
__global__ void k_cells_commit(uint32_t* ctl, uint32_t steps) {
    uint32_t lo = ctl[CTL_DONE_LO] + steps;
    if (lo < ctl[CTL_DONE_LO]) ctl[CTL_DONE_HI] += 1u;
    ctl[CTL_DONE_LO] = lo;
    ctl[CTL_CLOCK] += steps;
}

__global__ void k_flip_pp_pm(uint32_t* ctl) {
    ctl[CTL_PPOS] ^= 1u;
    ctl[CTL_PMETA] ^= 1u;
}

// ---------------------------------------------------------------------------------------------
// The leapfrog with its boundaries: Andersen bath, walls (hard, soft, heat bath), ledger
// (IntEngine3::step, int_engine3.rs:138-206; three.rs:232-265)
// ---------------------------------------------------------------------------------------------

// v += kick + g, the bath, x += v, then the boundaries.  Soft-wall impulses (wf, the force of
// each soft wall on this particle) go to the ledger as force x dt.  BOOK = false compiles the
// plain periodic / hard-wall leapfrog only (no ledger, no stochastic boundaries), so the hot
// kernels of systems without them carry none of that code's registers.
template <int D, bool BOOK>
__device__ __forceinline__ void advance(const Params<D>& P, const Bufs<D>& B, uint32_t i, uint32_t pm, uint32_t si, const uint32_t k[D], const double xi[D],
                                        uint32_t clock, const float* wf, double x[D], int32_t v[D]) {
    const bool pin = (P.pinned >> si) & 1u;
#pragma unroll
    for (int a = 0; a < D; a++) v[a] = (int32_t)((uint32_t)B.vel[pm][a][i] + k[a] + (pin ? 0u : (uint32_t)P.gravity[a]));
    if (!BOOK || !P.ledger) {
#pragma unroll
        for (int a = 0; a < D; a++) x[a] = xi[a] + (double)v[a];
        bound<D>(P, x, v);
        return;
    }
    const uint32_t id = B.ids[pm][i];
#pragma unroll
    for (int w = 0; w < 2 * D; w++)
        if (wf[w] != 0.0f) ledger_add(B.ledger, w, 1, (double)wf[w] * P.dt);
    if (P.bath) bath_redraw<D>(P, B.ledger, id, clock, si, v);
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = xi[a] + (double)v[a];
    bound_booked<D>(P, B.ledger, x, v, id, clock, si);
}

// ---------------------------------------------------------------------------------------------
// Rebuild chain
// ---------------------------------------------------------------------------------------------

template <int D>
__global__ __launch_bounds__(TPB) void k_keys(Params<D> P, Bufs<D> B, uint32_t* keys, uint32_t* vals, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    uint32_t pp = B.ctl[CTL_PPOS];
    double x[D];
    for (int a = 0; a < D; a++) x[a] = B.pos[pp][a][i];
    uint32_t c[D];
    cell_of<D>(P, x, c);
    keys[i] = morton<D>(c);
    vals[i] = i;
}

// Main bin of a position: row (z cell * cells along y + y cell), fine x cell.
__device__ __forceinline__ void row_bin(const Params<3>& P, const double x[3], uint32_t& row, uint32_t& xf) {
    uint32_t c[3];
    cell_of<3>(P, x, c);
    row = c[2] * P.nc[1] + c[1];
    const double f = x[0] * P.rows_xscale;
    const uint32_t v = f > 0.0 ? (uint32_t)f : 0u;
    xf = v < P.rows_nxf - 1u ? v : P.rows_nxf - 1u;
}

__global__ __launch_bounds__(TPB) void k_row_count(Params<3> P, Bufs<3> B, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS] ^ 1u;
    double x[3];
#pragma unroll
    for (int a = 0; a < 3; a++) x[a] = B.pos[pp][a][k];
    uint32_t row, xf;
    row_bin(P, x, row, xf);
    B.row_rank[k] = atomicAdd(&B.row_cnt[row * P.rows_nxf + xf], 1u);
}

// The rebuild's reorder on the row-build path: a counting sort of the Morton cell keys (one digit
// of all the key's bits), in place of the radix sort's passes.  Particles of one cell land in
// the order their atomic ranks fall; no result depends on the order within a cell (kicks are
// integer sums over the same pairs).
// The particles arrive in the previous cell order, so a wave's lanes mostly hold runs of equal
// keys: each run takes its ranks with one atomic (made by the run's first lane) instead of one
// per particle contending for the same counter.
template <int D>
__global__ __launch_bounds__(TPB) void k_keys_count(Params<D> P, Bufs<D> B, uint32_t* keys, uint32_t* ranks, uint32_t* count, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    const bool live = i < P.n;
    uint32_t key = 0xffffffffu;  // lanes past the end form a run of their own (keys are below 2^30)
    if (live) {
        const uint32_t pp = B.ctl[CTL_PPOS];
        double x[D];
#pragma unroll
        for (int a = 0; a < D; a++) x[a] = B.pos[pp][a][i];
        uint32_t c[D];
        cell_of<D>(P, x, c);
        key = morton<D>(c);
    }
    const uint32_t lane = __lane_id();
    const uint32_t prev = __shfl_up(key, 1, 64);
    const uint64_t heads = ballot64(lane == 0u || prev != key);
    const uint64_t upto = lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull);  // lanes 0 .. lane
    const uint32_t start = 63u - (uint32_t)__clzll(heads & upto);
    const uint64_t after = heads & ~upto;
    const uint32_t end = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
    uint32_t base = 0u;
    if (live && lane == start) base = atomicAdd(&count[key], end - start);
    base = __shfl(base, (int)start, 64);
    if (!live) return;
    keys[i] = key;
    ranks[i] = base + (lane - start);
}

// perm[first slot of the key + rank] = the particle's current slot
__global__ __launch_bounds__(TPB) void k_keys_place(uint32_t n, const uint32_t* keys, const uint32_t* ranks, const uint32_t* first, uint32_t* perm, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    perm[first[keys[i]] + ranks[i]] = i;
}

__global__ void k_ranges(const uint32_t* keys, uint32_t n, uint32_t* start, uint32_t* end, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= n) return;
    uint32_t key = keys[k];
    if (k == 0 || keys[k - 1] != key) start[key] = k;
    if (k == n - 1 || keys[k + 1] != key) end[key] = k + 1;
}

// Reorder storage by the sorted permutation into the other buffers; positions also become the
// list's reference positions.
template <int D>
__global__ __launch_bounds__(TPB) void k_gather(Params<D> P, Bufs<D> B, const uint32_t* perm, int write_ref, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= P.n) return;
    uint32_t pp = B.ctl[CTL_PPOS], pm = B.ctl[CTL_PMETA];
    uint32_t i = perm[k];
    double xs[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double x = B.pos[pp][a][i];
        xs[a] = x;
        B.pos[pp ^ 1][a][k] = x;
        if (write_ref) B.ref[a][k] = x;
        B.vel[pm ^ 1][a][k] = B.vel[pm][a][i];
    }
    // the packed copy holds sixteenths (the step kernel's gathers take them as they are)
    if constexpr (D == 3) reinterpret_cast<double4*>(B.aos[pp ^ 1])[k] = make_double4(xs[0] * 0.0625, xs[1] * 0.0625, xs[2] * 0.0625, 0.0);
    else reinterpret_cast<double2*>(B.aos[pp ^ 1])[k] = make_double2(xs[0] * 0.0625, xs[1] * 0.0625);
    uint32_t c[D];
    cell_of<D>(P, xs, c);
    float r[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) r[a] = (float)(xs[a] * 0x1p-32 - (double)c[a] * P.cell_w64[a]);
    B.rel[k] = make_float4(r[0], r[1], r[2], 0.f);
    B.species[pm ^ 1][k] = B.species[pm][i];
    B.ids[pm ^ 1][k] = B.ids[pm][i];
}

// Full Verlet list within each species pair's (rc + skin)(1 + 1e-4) (md3.wgsl build_list), over
// the 3^D cells around the particle's, in the freshly gathered (not yet committed) buffers.
// Offsets of the 3^D stencil that revisit a cell: on an axis with a single cell only offset 0.
template <int D, class PT>
__device__ __forceinline__ bool stencil_dup(const PT& P, int o) {
    int off[3] = {o % 3 - 1, o / 3 % 3 - 1, o / 9 - 1};
    bool dup = false;
#pragma unroll
    for (int a = 0; a < D; a++) dup |= (P.nc[a] == 1u && off[a] != 0);
    return dup;
}

template <int D, bool MULTI>
__global__ __launch_bounds__(TPB) void k_build_list(Params<D> P, Bufs<D> B, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    uint32_t i = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
    if (i >= P.n) return;
    uint32_t pp = B.ctl[CTL_PPOS] ^ 1u, pm = B.ctl[CTL_PMETA] ^ 1u;
    const double* px[D];
    for (int a = 0; a < D; a++) px[a] = B.pos[pp][a];
    const uint32_t* sp = B.species[pm];
    double xi[D];
    for (int a = 0; a < D; a++) xi[a] = px[a][i];
    uint32_t si = MULTI ? sp[i] : 0u;
    const float list_r2 = B.pt->list_r2[0];
    uint32_t c[D];
    cell_of<D>(P, xi, c);
    uint32_t count = 0;
    const uint32_t cap = P.cap, n = P.n;
    for (int o = 0; o < stencil_size(D); o++) {
        if (stencil_dup<D>(P, o)) continue;
        uint32_t cc[D];
        stencil_cell<D>(P, c, o, cc);
        uint32_t key = morton<D>(cc);
        uint32_t s = B.cell_start[key], e = B.cell_end[key];
        for (uint32_t j = s; j < e; j++) {
            if (j == i) continue;
            float f[D];
            bool ok = true;
#pragma unroll
            for (int a = 0; a < D; a++) ok &= sep_f(separation<D>(P, xi[a], px[a][j], a), f[a]);
            float r2 = f[0] * f[0] + f[1] * f[1];
            if constexpr (D == 3) r2 = r2 + f[2] * f[2];
            float lr2 = MULTI ? B.pt->list_r2[si * 4 + sp[j]] : list_r2;
            if (ok && r2 < lr2) {
                if (count < cap) B.nbr[(size_t)count * n + i] = j;
                count++;
            }
        }
    }
    if (count > cap) {
        atomicOr(&B.ctl[CTL_OVERFLOW], 1u);
    }
    if (count > *(volatile const uint32_t*)&B.ctl[CTL_MAXCOUNT]) atomicMax(&B.ctl[CTL_MAXCOUNT], count);
    B.nbr_count[i] = count < cap ? count : cap;
}

// ---------------------------------------------------------------------------------------------
// The step: forces over the list, leapfrog, half-skin check (one launch)
// ---------------------------------------------------------------------------------------------

template <int D, bool MULTI, bool BOOK>
__global__ __launch_bounds__(TPB) void k_step_list(Params<D> P, Bufs<D> B) {
    const uint32_t* ctl = B.ctl;
    if (!step_runs(ctl)) return;
    uint32_t i = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = ctl[CTL_PPOS], pm = ctl[CTL_PMETA];
    const double* px[D];
    for (int a = 0; a < D; a++) px[a] = B.pos[pp][a];
    const uint32_t* sp = B.species[pm];
    double xi[D];
    for (int a = 0; a < D; a++) xi[a] = px[a][i];
    uint32_t si = MULTI ? sp[i] : 0u;
    const PairTable* __restrict__ pt = B.pt;
    float c = MULTI ? pt->kick_c[si] : P.kick_c[0];
    const float rc2_0 = pt->rc2[0], sigma2_0 = pt->sigma2[0], six_eps4_0 = pt->six_eps4[0], r2f_0 = pt->r2_fast[0];
    uint32_t k[D];
    for (int a = 0; a < D; a++) k[a] = 0u;
    const uint32_t m = B.nbr_count[i];
    const uint32_t* nb = B.nbr + i;
    const size_t stride = P.n;
    uint32_t j_next = m ? nb[0] : 0u;
    for (uint32_t e = 0; e < m; e++) {
        const uint32_t j = j_next;
        if (e + 1 < m) j_next = nb[(size_t)(e + 1) * stride];
        double xj[D];
#pragma unroll
        for (int a = 0; a < D; a++) xj[a] = px[a][j];
        float f[D];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < D; a++) ok &= sep_f(separation<D>(P, xi[a], xj[a], a), f[a]);
        if constexpr (MULTI) {
            uint32_t p = si * 4 + sp[j];
            if (ok) pair_kick<D>(f, pt->rc2[p], pt->sigma2[p], pt->six_eps4[p], pt->r2_fast[p], c, k);
        } else {
            if (ok) pair_kick<D>(f, rc2_0, sigma2_0, six_eps4_0, r2f_0, c, k);
        }
    }
    float wf[2 * D] = {};
    if (P.soft_walls) soft_walls<D>(P, xi, c, k, nullptr, wf);
    // leapfrog: v += kick + g, x += v, boundaries
    int32_t v[D];
    double x[D];
    advance<D, BOOK>(P, B, i, pm, si, k, xi, ctl[CTL_CLOCK], wf, x, v);
    float r2 = 0.0f;
    bool far = false;
#pragma unroll
    for (int a = 0; a < D; a++) {
        B.pos[pp ^ 1][a][i] = x[a];
        B.vel[pm][a][i] = v[a];
        float d;
        far |= !sep_f(separation<D>(P, B.ref[a][i], x[a], a), d);
        r2 += d * d;
    }
    bool stale = far || r2 >= P.half_skin2;
    uint64_t bal = ballot64(stale);
    if (bal != 0ull && __lane_id() == (uint32_t)(__ffsll((long long)bal) - 1) && *(volatile const uint32_t*)&B.ctl[CTL_STALE_NEXT] == 0u)
        atomicOr(&B.ctl[CTL_STALE_NEXT], 1u);
}


// ---------------------------------------------------------------------------------------------
// Half-list variant
// ---------------------------------------------------------------------------------------------

// List build from f32 cell-relative coordinates: d = (r_j + offset of j's cell) - r_i, within
// ~1e-6 sigma of the exact separation, tested against the same (rc + skin)(1 + 1e-4) radius, so
// every pair the exact test admits is admitted (the 1e-4 margin is ~3e-4 sigma).  Walled axes do
// not wrap the stencil.  In-block pairs go to their lower-index end only (the step kernel deals
// chunks out evenly, so how pairs divide between particles costs it nothing); pairs across
// blocks go to both ends.  The list is a stream of four-entry chunks (16 B).  Entries are
// particle indices below 2^28; the top four bits of each word carry the chunk's tag
// (tag_chunk): the owner's index in its block of P.block particles, the entry count and whether
// the entries are in that block.  Every 64-particle wave fills a region of 64 cap/4 slots from
// its start, chunks in the order they fill, and records its chunk count.  The candidate and hit
// loops run in step across the wave (a lane with nothing left idles), so the stores happen
// where every lane is present and a wave-uniform count hands out the slots.
template <int D, bool MULTI, bool WRAP>
__global__ __launch_bounds__(TPB) void k_build_half(Params<D> P, Bufs<D> B, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS] ^ 1u, pm = B.ctl[CTL_PMETA] ^ 1u;
    const uint32_t* sp = B.species[pm];
    const float4* __restrict__ rel = B.rel;
    double xi[D];
#pragma unroll
    for (int a = 0; a < D; a++) xi[a] = B.pos[pp][a][i];
    const float4 me = rel[i];
    const float ri[3] = {me.x, me.y, me.z};
    const uint32_t si = MULTI ? sp[i] : 0u;
    const float list_r2 = B.pt->list_r2[0];
    uint32_t c[D];
    cell_of<D>(P, xi, c);
    const uint32_t cap4 = P.cap / 4u;
    const uint32_t blo = i & ~(P.block - 1u);  // the block is a power of two
    const uint32_t wave = i >> 6, own = i & (P.block - 1u);
    const uint32_t capw = 64u * cap4;
    uint4* chunks = reinterpret_cast<uint4*>(B.nbr) + (size_t)wave * capw;
    const uint64_t lt = lanemask_lt();
    // A lane fills one in-block and one cross-block chunk in registers; full chunks are stored
    // whole (16 B) in the wave's region, in-block chunks from its start and cross-block ones from
    // its end down.  The fills change only where every lane is present.
    uint32_t fill_in = 0u, fill_out = 0u;
    auto store = [&](bool want, const uint4& q, bool in) {
        const uint64_t bal = ballot64(want);
        if (bal == 0ull) return;
        if (want) {
            const uint32_t k = (in ? fill_in : fill_out) + (uint32_t)__popcll(bal & lt), other = in ? fill_out : fill_in;
            if (k + other < capw) chunks[in ? k : capw - 1u - k] = q;
            else atomicOr(&B.ctl[CTL_OVERFLOW], 1u);
        }
        if (in) fill_in += (uint32_t)__popcll(bal);
        else fill_out += (uint32_t)__popcll(bal);
    };
    // the open in-block chunk (eight offsets from the block's first slot) and cross-block chunk
    uint4 fa0 = make_uint4(0u, 0u, 0u, 0u), fa1 = make_uint4(0u, 0u, 0u, 0u), fb = make_uint4(0u, 0u, 0u, 0u);
    auto inblock_chunk = [&](uint32_t entries) {
        const uint32_t e[8] = {fa0.x, fa0.y, fa0.z, fa0.w, fa1.x, fa1.y, fa1.z, fa1.w};
        return pack_inblock(e, own, entries);
    };
    // single-cell periodic axes: the nearest image of the one cell (cell_w is the box length)
    float wl[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) wl[a] = (P.nc[a] == 1u && ((P.periodic >> a) & 1u)) ? P.cell_w[a] : 0.f;
    uint32_t na = 0, nb = 0;
    // Hit masks wait in a three-entry queue per lane (mask, first candidate) and are visited
    // one hit per step of a wave-wide loop, run only when some lane's queue is full and at the
    // end: the loop's length follows the busiest lane's hits over many cells, not the sum over
    // cells of each cell's busiest lane.
    uint32_t m0 = 0u, m1 = 0u, m2 = 0u, b0 = 0u, b1 = 0u, b2 = 0u, qn = 0u;
    auto visit = [&]() {
        bool full = false, in = false;
        if (qn != 0u) {
            const uint32_t b = __builtin_ctz(m0);
            m0 &= m0 - 1u;
            const uint32_t j = b0 + b;
            in = j - blo < P.block;
            if (in) {
                if ((na & 7u) < 4u) put4(fa0, na, j - blo);
                else put4(fa1, na, j - blo);
                na++;
                full = (na & 7u) == 0u;
            } else {
                put4(fb, nb, j);
                nb++;
                full = (nb & 3u) == 0u;
            }
            if (m0 == 0u) {
                m0 = m1; b0 = b1; m1 = m2; b1 = b2; m2 = 0u;
                qn--;
            }
        }
        const bool fin = full && in, fout = full && !in;
        store(fin, fin ? inblock_chunk(8u) : make_uint4(0u, 0u, 0u, 0u), true);
        store(fout, tag_chunk(fb, own, 4u, false), false);
    };
    for (int o = 0; o < stencil_size(D); o++) {
        const int off[3] = {o % 3 - 1, o / 3 % 3 - 1, o / 9 - 1};
        uint32_t cc[D];
        float shift[3] = {0.f, 0.f, 0.f};
        bool skip = false;
#pragma unroll
        for (int a = 0; a < D; a++) {
            int v = (int)c[a] + off[a];
            if (P.nc[a] == 1u && off[a] != 0) skip = true;
            if (v < 0 || v >= (int)P.nc[a]) {
                if (!((P.periodic >> a) & 1u)) skip = true;
                v = v < 0 ? v + (int)P.nc[a] : v - (int)P.nc[a];
            }
            cc[a] = (uint32_t)v;
            shift[a] = (float)off[a] * P.cell_w[a] - ri[a];
        }
        uint32_t base = 0u, e0 = 0u;
        if (!skip) {
            const uint32_t key = morton<D>(cc);
            const uint32_t s0 = B.cell_start[key];
            e0 = B.cell_end[key];
            // An in-block pair belongs to its lower-index end, so in-block candidates up to i
            // are not this end's: a cell that begins inside the block is scanned from past i.
            base = s0 >= blo ? max(s0, i + 1u) : s0;
        }
        const float sx = shift[0], sy = shift[1], sz = shift[2];
        // Candidates in groups of up to 32: a distance pass sets one bit per candidate within
        // the list radius, bit arithmetic drops the in-block ones up to i (i itself included),
        // and a nonzero mask joins the queue.
        for (;; base += 32u) {
            const bool act = base < e0;
            if (ballot64(act) == 0ull) break;
            uint32_t mask = 0u;
            if (act) {
                const uint32_t cnt = min(32u, e0 - base);
                // Candidates four at a time, unrolled (addresses by immediate offset, the bit by
                // a select); a lane past its cell's end tests the rel entries that follow, which
                // exist (the array is padded by 32) and whose bits are masked off after.
                const float4* __restrict__ rq = rel + base;
                auto near = [&](uint32_t u) -> bool {
                    const float4 q = rq[u];
                    float dx = q.x + sx, dy = q.y + sy, dz = D == 3 ? q.z + sz : 0.f;
                    if constexpr (WRAP) {
                        if (wl[0] != 0.f) dx = dx > 0.5f * wl[0] ? dx - wl[0] : (dx < -0.5f * wl[0] ? dx + wl[0] : dx);
                        if (wl[1] != 0.f) dy = dy > 0.5f * wl[1] ? dy - wl[1] : (dy < -0.5f * wl[1] ? dy + wl[1] : dy);
                        if (D == 3 && wl[2] != 0.f) dz = dz > 0.5f * wl[2] ? dz - wl[2] : (dz < -0.5f * wl[2] ? dz + wl[2] : dz);
                    }
                    float r2 = dx * dx;
                    r2 = __builtin_fmaf(dy, dy, r2);
                    if (D == 3) r2 = __builtin_fmaf(dz, dz, r2);
                    const float lr2 = MULTI ? B.pt->list_r2[si * 4 + sp[min(base + u, e0 - 1u)]] : list_r2;
                    return r2 < lr2;
                };
#pragma unroll 1
                for (uint32_t u0 = 0; u0 < 32u; u0 += 4u) {
                    if (u0 >= cnt) break;
#pragma unroll
                    for (uint32_t k = 0; k < 4u; k++) mask |= near(u0 + k) ? 1u << (u0 + k) : 0u;
                }
                mask &= bits_below((int)cnt);
                // j = base + u: gt: j > i; inb: j in i's block
                const uint32_t gt = ~bits_below((int)i + 1 - (int)base);
                const uint32_t inb = bits_below((int)(blo + P.block) - (int)base) & ~bits_below((int)blo - (int)base);
                mask &= (inb & gt) | ~inb;
            }
            const bool push = mask != 0u;
            while (ballot64(push && qn == 3u) != 0ull) visit();
            if (push) {
                if (qn == 0u) { m0 = mask; b0 = base; }
                else if (qn == 1u) { m1 = mask; b1 = base; }
                else { m2 = mask; b2 = base; }
                qn++;
            }
        }
    }
    while (ballot64(qn != 0u) != 0ull) visit();
    // the partial last chunks; a list longer than the capacity is flagged and regrown
    const bool pin = (na & 7u) != 0u, pout = (nb & 3u) != 0u;
    store(pin, pin ? inblock_chunk(na & 7u) : make_uint4(0u, 0u, 0u, 0u), true);
    store(pout, tag_chunk(fb, own, pout ? nb & 3u : 1u, false), false);
    const uint32_t ca = (na + 7u) / 8u, cb = (nb + 3u) / 4u;
    if (ca + cb > cap4) atomicOr(&B.ctl[CTL_OVERFLOW], 1u);
    if (4u * (ca + cb) > *(volatile const uint32_t*)&B.ctl[CTL_MAXCOUNT]) atomicMax(&B.ctl[CTL_MAXCOUNT], 4u * (ca + cb));
    if (__lane_id() == 0u) {
        const uint32_t in = min(fill_in, capw);
        B.wave_chunks[2u * wave] = in;
        B.wave_chunks[2u * wave + 1u] = min(fill_out, capw - in);
    }
}

// ---------------------------------------------------------------------------------------------
// Row-ordered half-list build (3D)
// ---------------------------------------------------------------------------------------------
// After the reorder, a second copy of the particles is ordered by row and fine x bin: a row is
// a (y, z) cell, and along x each row is binned into fine cells (counting sort: count,
// exclusive scan, scatter; the order within a fine cell is arbitrary, which no result depends
// on, since list order only changes the order of integer sums).  On a periodic x axis every row
// also carries ghost copies of its first and last few bins beyond its ends (shifted by the box),
// so that the candidates whose x lies within the list sphere's chord over a row's (y, z)
// rectangle are always one contiguous range of the copy.  The sphere around a particle meets at
// most the nine rows around its own: about 257 candidates per particle at liquid density against
// about 513 in the 27 cells of the stencil build.  Candidates are tested as in the stencil build
// (f32 separations within ~1e-5 sigma of the exact one against the (rc + skin) (1 + 1e-4)
// radius, a margin of ~3e-4 sigma), with x absolute (boxes up to 512 sigma along x) and y, z
// relative to the cell's corner.  The chunk stream written is the stencil build's: the same
// pairs, tags and wave regions, in another order.


// Counts of the extended bins (ghost bins take their source bin's count), for the scan; the
// entry past the last is zeroed, so the scan leaves the total there.
__global__ __launch_bounds__(TPB) void k_row_ext(Params<3> P, Bufs<3> B, uint32_t nkeys, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k > nkeys) return;
    if (k == nkeys) {
        B.row_start[k] = 0u;
        return;
    }
    const uint32_t row = k / P.rows_ne, e = k - row * P.rows_ne;
    int b = (int)e - (int)P.rows_g;
    if (b < 0) b += (int)P.rows_nxf;
    else if (b >= (int)P.rows_nxf) b -= (int)P.rows_nxf;
    B.row_start[k] = B.row_cnt[row * P.rows_nxf + (uint32_t)b];
}

__global__ __launch_bounds__(TPB) void k_row_scatter(Params<3> P, Bufs<3> B, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS] ^ 1u, pm = B.ctl[CTL_PMETA] ^ 1u;
    double x[3];
#pragma unroll
    for (int a = 0; a < 3; a++) x[a] = B.pos[pp][a][k];
    uint32_t row, xf;
    row_bin(P, x, row, xf);
    const float4 r = B.rel[k];
    const float ry = r.y, rz = r.z;
    const uint32_t rank = B.row_rank[k], base = row * P.rows_ne, g = P.rows_g, nxf = P.rows_nxf;
    const uint32_t sp = P.nspecies > 1 ? B.species[pm][k] << 30 : 0u;
    const float w = __uint_as_float(k | sp);
    B.items[B.row_start[base + g + xf] + rank] = make_float4((float)(x[0] * 0x1p-32), ry, rz, w);
    if (xf < g)  // a ghost past the row's end
        B.items[B.row_start[base + g + nxf + xf] + rank] = make_float4((float)((x[0] + P.boxf[0]) * 0x1p-32), ry, rz, w);
    if (xf + g >= nxf)  // a ghost before the row's start
        B.items[B.row_start[base + xf + g - nxf] + rank] = make_float4((float)((x[0] - P.boxf[0]) * 0x1p-32), ry, rz, w);
}

template <bool MULTI>
__global__ __launch_bounds__(TPB) void k_build_rows(Params<3> P, Bufs<3> B, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS] ^ 1u, pm = B.ctl[CTL_PMETA] ^ 1u;
    uint32_t c[3];
    float Xi;
    {
        double xi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) xi[a] = B.pos[pp][a][i];
        cell_of<3>(P, xi, c);
        Xi = (float)(xi[0] * 0x1p-32);
    }
    float mey, mez;  // y and z relative to the cell corner
    {
        const float4 me = B.rel[i];
        mey = me.y;
        mez = me.z;
    }
    float lr2[4];
    if constexpr (MULTI) {
        const uint32_t si = B.species[pm][i];
#pragma unroll
        for (uint32_t s = 0; s < 4; s++) lr2[s] = B.pt->list_r2[si * 4 + s];
    } else {
        lr2[0] = B.pt->list_r2[0];
    }
    const float reach2 = P.rows_reach * P.rows_reach;
    const float4* __restrict__ items = B.items;
    const uint32_t cap4 = P.cap / 4u;
    // the block (a power of two, at least a wave) and the wave are the same for every lane
    const uint32_t blo = __builtin_amdgcn_readfirstlane(i & ~(P.block - 1u));
    const uint32_t wave = __builtin_amdgcn_readfirstlane(i >> 6), own = i & (P.block - 1u);
    const uint32_t capw = 64u * cap4;
    uint4* chunks = reinterpret_cast<uint4*>(B.nbr) + (size_t)wave * capw;
    // Full chunks are stored whole (16 B) in the wave's region: in-block chunks from its start,
    // cross-block chunks from its end down.  The fills change only where every lane is present.
    uint32_t fill_in = 0u, fill_out = 0u;
    auto store = [&](bool want, const uint4& q, bool in) {
        const uint64_t bal = ballot64(want);
        if (bal == 0ull) return;
        if (want) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            const uint32_t k = (in ? fill_in : fill_out) + below, other = in ? fill_out : fill_in;
            if (k + other < capw) chunks[in ? k : capw - 1u - k] = q;
            else atomicOr(&B.ctl[CTL_OVERFLOW], 1u);
        }
        if (in) fill_in += (uint32_t)__popcll(bal);
        else fill_out += (uint32_t)__popcll(bal);
    };
    // Per lane, in LDS: the open in-block chunk (up to eight offsets from the block's first
    // slot, entry s at ci[s][lane]) and cross-block chunk (four slots, co[s][lane]), and a ring
    // of up to four hit masks, each with the copy index of its bit 0 (bit b stands for that index
    // minus b) and its kind in bit 31.  A visit moves as many hits of the front mask into that
    // kind's open chunk as it has room for (up to four) and stores the chunk when it is full.
    // Visits run when some lane's ring has no room for a group's two masks, and at the end.
    // (Held in LDS, the ring and chunks leave registers to the candidate loop.)
    __shared__ uint32_t ci[8][TPB], co[4][TPB];
    __shared__ uint32_t qm[4][TPB], qt[4][TPB];
    const uint32_t t = threadIdx.x;
    uint32_t na = 0u, nb = 0u, head = 0u, qn = 0u;
    auto inblock_chunk = [&](uint32_t entries) {
        uint32_t e[8];
#pragma unroll
        for (int s2 = 0; s2 < 8; s2++) e[s2] = ci[s2][t];
        return pack_inblock(e, own, entries);
    };
    auto cross_chunk = [&](uint32_t entries) { return tag_chunk(make_uint4(co[0][t], co[1][t], co[2][t], co[3][t]), own, entries, false); };
    auto visit = [&]() {
        bool full = false, in = false;
        if (qn != 0u) {
            const uint32_t h = head & 3u;
            uint32_t m = qm[h][t];
            const uint32_t tag = qt[h][t];
            in = (tag >> 31) != 0u;
            const uint32_t b1 = tag & 0x7fffffffu;  // the copy index of bit 0
            const uint32_t size = in ? 8u : 4u, n = in ? na : nb, r = n & (size - 1u);
            uint32_t jv[4];
            bool take[4];
#pragma unroll
            for (uint32_t s2 = 0; s2 < 4u; s2++) {
                take[s2] = r + s2 < size && m != 0u;
                const uint32_t at = take[s2] ? (uint32_t)__builtin_ctz(m) : 0u;
                m = take[s2] ? m & (m - 1u) : m;
                jv[s2] = __float_as_uint(items[b1 - at].w) & 0x0fffffffu;
            }
#pragma unroll
            for (uint32_t s2 = 0; s2 < 4u; s2++) {
                if (!take[s2]) continue;
                if (in) ci[r + s2][t] = jv[s2] - blo;
                else co[r + s2][t] = jv[s2];
            }
            const uint32_t taken = (uint32_t)take[0] + (uint32_t)take[1] + (uint32_t)take[2] + (uint32_t)take[3];
            const uint32_t n2 = n + taken;
            full = (n2 & (size - 1u)) == 0u;  // at least one was taken, so the chunk closed
            if (in) na = n2;
            else nb = n2;
            if (m == 0u) {
                head++;
                qn--;
            } else {
                qm[h][t] = m;
            }
        }
        // every lane takes part in the stores' ballots, so the fills stay the same across the wave
        const bool fin = full && in, fout = full && !in;
        store(fin, fin ? inblock_chunk(8u) : make_uint4(0u, 0u, 0u, 0u), true);
        store(fout, fout ? cross_chunk(4u) : make_uint4(0u, 0u, 0u, 0u), false);
    };
    auto push = [&](uint32_t m, uint32_t tag) {
        if (m == 0u) return;
        const uint32_t h = (head + qn) & 3u;
        qm[h][t] = m;
        qt[h][t] = tag;
        qn++;
    };
    // the chord's bins are clamped to the row's extended bins (its ghosts on a periodic x axis)
    const int glo = -(int)P.rows_g, ghi = (int)(P.rows_nxf + P.rows_g) - 1;
    const uint32_t ny = P.nc[1], nz = P.nc[2];
    const float wy = P.cell_w[1], wz = P.cell_w[2], sx = -Xi;
    for (int o = 0; o < 9; o++) {
        const int oy = o % 3 - 1, oz = o / 3 - 1;
        int yy = (int)c[1] + oy, zz = (int)c[2] + oz;
        bool skip = false;
        if (yy < 0 || yy >= (int)ny) {
            if (!((P.periodic >> 1) & 1u)) skip = true;
            yy = yy < 0 ? yy + (int)ny : yy - (int)ny;
        }
        if (zz < 0 || zz >= (int)nz) {
            if (!((P.periodic >> 2) & 1u)) skip = true;
            zz = zz < 0 ? zz + (int)nz : zz - (int)nz;
        }
        // distance from the particle to the row's (y, z) rectangle, and the chord along x
        const float dyl = oy == 0 ? 0.f : fmaxf(0.f, oy > 0 ? wy - mey : mey);
        const float dzl = oz == 0 ? 0.f : fmaxf(0.f, oz > 0 ? wz - mez : mez);
        const float d2 = dyl * dyl + dzl * dzl;
        skip |= !(d2 < reach2);
        const float ext = skip ? 0.f : __builtin_sqrtf(reach2 - d2);
        const float sy = (float)oy * wy - mey, sz = (float)oz * wz - mez;
        int a = (int)__builtin_floorf((Xi - ext) * P.rows_inv_wx);
        int b = (int)__builtin_floorf((Xi + ext) * P.rows_inv_wx);
        a = a > glo ? a : glo;
        b = b < ghi ? b : ghi;
        uint32_t s0 = 0u, e0 = 0u;
        if (!skip && a <= b) {
            const uint32_t rowbase = ((uint32_t)zz * ny + (uint32_t)yy) * P.rows_ne + P.rows_g;
            s0 = B.row_start[rowbase + (uint32_t)a];
            e0 = B.row_start[rowbase + (uint32_t)b + 1u];
        }
        {
            // Candidates in groups of up to 32, tested without a branch (each candidate is one
            // 16-byte load) into two masks: in-block partners of higher index than i (an
            // in-block pair is kept at its lower-index end only) and partners outside i's block,
            // both within the list radius.
            for (uint32_t base = s0;; base += 32u) {
                const bool act = base < e0;
                if (ballot64(act) == 0ull) break;
                // Hits are shifted in as they are tested: candidate base + u ends at bit
                // n - 1 - u of its mask, n being the number of candidates tested.
                uint32_t m_in = 0u, m_out = 0u, last = base;
                if (act) {
                    const uint32_t cnt = min(32u, e0 - base);
                    // four at a time; a lane past its range's end tests the copy's entries that
                    // follow (the copy is padded by 32), whose bits are masked off after
                    const float4* __restrict__ rq = items + base;
#pragma unroll 1
                    for (uint32_t u0 = 0; u0 < 32u; u0 += 4u) {
                        if (u0 >= cnt) break;
#pragma unroll
                        for (uint32_t k = 0; k < 4u; k++) {
                            const float4 q = rq[u0 + k];
                            const float dx = q.x + sx, dy = q.y + sy, dz = q.z + sz;
                            float r2 = dx * dx;
                            r2 = __builtin_fmaf(dy, dy, r2);
                            r2 = __builtin_fmaf(dz, dz, r2);
                            const uint32_t w = __float_as_uint(q.w);
                            uint32_t j;
                            float lr;
                            if constexpr (MULTI) {
                                j = w & 0x0fffffffu;
                                const uint32_t sj = w >> 30;
                                lr = sj == 0u ? lr2[0] : (sj == 1u ? lr2[1] : (sj == 2u ? lr2[2] : lr2[3]));
                            } else {
                                j = w;  // one species: the copy holds the slot alone
                                lr = lr2[0];
                            }
                            // the three tests as wave masks, combined on the scalar unit
                            const uint64_t near = ballot64(r2 < lr), inb = ballot64(j - blo < P.block), gt = ballot64(j > i);
                            m_in = shift_in(m_in, near & inb & gt);
                            m_out = shift_in(m_out, near & ~inb);
                        }
                    }
                    const uint32_t n = (cnt + 3u) & ~3u;
                    const uint32_t live = ~bits_below((int)(n - cnt));
                    m_in &= live;
                    m_out &= live;
                    last = base + n - 1u;
                }
                const uint32_t need = (m_in != 0u ? 1u : 0u) + (m_out != 0u ? 1u : 0u);
                while (ballot64(qn + need > 4u) != 0ull) visit();
                push(m_in, last | 0x80000000u);
                push(m_out, last);
            }
        }
    }
    while (ballot64(qn != 0u) != 0ull) visit();
    // the partial last chunks; a list longer than the capacity is flagged and regrown
    const bool pin = (na & 7u) != 0u, pout = (nb & 3u) != 0u;
    store(pin, pin ? inblock_chunk(na & 7u) : make_uint4(0u, 0u, 0u, 0u), true);
    store(pout, pout ? cross_chunk(nb & 3u) : make_uint4(0u, 0u, 0u, 0u), false);
    const uint32_t ca = (na + 7u) / 8u, cb = (nb + 3u) / 4u;
    if (ca + cb > cap4) atomicOr(&B.ctl[CTL_OVERFLOW], 1u);
    if (4u * (ca + cb) > *(volatile const uint32_t*)&B.ctl[CTL_MAXCOUNT]) atomicMax(&B.ctl[CTL_MAXCOUNT], 4u * (ca + cb));
    if (__lane_id() == 0u) {
        const uint32_t in = min(fill_in, capw);
        B.wave_chunks[2u * wave] = in;
        B.wave_chunks[2u * wave + 1u] = min(fill_out, capw - in);
    }
}

// A block of BS particles runs on min(BS, 1024) threads (with BS = 2048 each thread advances
// particles t and t + 1024).  The block's list chunks (its build waves' regions, numbered in
// order) are dealt out evenly: lane t takes chunks t, t + threads, ... whichever particles own
// them, so a wave's lanes stay busy however unevenly the pairs divide between particles, and
// neighbouring lanes read neighbouring chunks.  Every kick goes to the block's LDS
// accumulators: the owner's after each chunk, an in-block partner's per pair.
template <int D, bool MULTI, int BS, bool BOOK>
__global__ __launch_bounds__(BS > 1024 ? 1024 : BS, BS == 2048 && !BOOK ? 8 : 1) void k_step_half(Params<D> P, Bufs<D> B) {
    constexpr uint32_t NT = BS > 1024 ? 1024u : (uint32_t)BS;
    constexpr uint32_t PER = (uint32_t)BS / NT;
    const uint32_t* ctl = B.ctl;
    if (!step_runs(ctl)) return;
    __shared__ uint32_t acc[D][BS];
    // The block's own positions, in sixteenths (x / 16, exact): in-block neighbours are read from
    // LDS, not gathered from memory.  The force loop works on sixteenths throughout.
    __shared__ double lpos[D][BS];
    const uint32_t t = threadIdx.x;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = blk * BS;
    const uint32_t pp = ctl[CTL_PPOS], pm = ctl[CTL_PMETA];
    const double* px[D];
    for (int a = 0; a < D; a++) px[a] = B.pos[pp][a];
    // Periodic axes on which the block needs no image: with every particle of the block at
    // least face16 from both faces (the bound on a listed partner's distance at step time: the
    // largest list radius plus the two particles' moves since the build), every partner's stored
    // position is its nearest image, and separation16's correction is exactly zero.
    __shared__ uint32_t edge;
    if (t == 0u) edge = 0u;
    __syncthreads();
    uint32_t near_face = 0u;
#pragma unroll
    for (uint32_t r = 0; r < PER; r++) {
        const uint32_t l = t + r * NT;
#pragma unroll
        for (int a = 0; a < D; a++) {
            const bool real = base + l < P.n;
            const double x16 = real ? px[a][base + l] * 0.0625 : 0.0;
            lpos[a][l] = x16;
            acc[a][l] = 0u;
            if (real && ((P.periodic >> a) & 1u) && !(x16 >= P.face16 && x16 <= P.box16[a] - P.face16)) near_face = 1u;
        }
    }
    if (ballot64(near_face != 0u) != 0ull && __lane_id() == 0u) atomicOr(&edge, 1u);
    // the in-block and cross-block chunk counts of the block's NW list-build waves, scanned by
    // the first wave
    constexpr uint32_t NW = (uint32_t)BS / 64u;
    __shared__ uint32_t wfirst[2][NW + 1];
    if (t < 64u) {
#pragma unroll
        for (uint32_t kind = 0; kind < 2u; kind++) {
            const uint32_t v = t < NW ? B.wave_chunks[2u * (blk * NW + t) + kind] : 0u;
            uint32_t x = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if (t >= (uint32_t)d) x += y;
            }
            if (t < NW) wfirst[kind][t] = x - v;
            if (t == NW - 1u) wfirst[kind][NW] = x;
        }
    }
    __syncthreads();
    const bool images = __builtin_amdgcn_readfirstlane(edge) != 0u;
    const uint32_t* sp = B.species[pm];
    const PairTable* __restrict__ pt = B.pt;
    {
        const float rc2_0 = pt->rc2[0], sigma2_0 = pt->sigma2[0], six_eps4_0 = pt->six_eps4[0], r2f_0 = pt->r2_fast[0];
        // The block's chunks of each kind, numbered 0 .. n-1 across its waves' regions in order:
        // in-block chunk g lies in region w (wfirst[0][w] <= g < wfirst[0][w + 1]) at slot
        // g - wfirst[0][w], cross-block chunk g at slot capw - 1 - (g - wfirst[1][w]).
        const size_t capw = 64u * (P.cap / 4u);
        const uint4* __restrict__ chunks = reinterpret_cast<const uint4*>(B.nbr) + (size_t)blk * NW * capw;
        // one pair: its separation (with or without the periodic image), its kick on the owner
        // (summed into k) and its kick on the partner (kj); false beyond the cutoff
        auto pair = [&](auto with_images, const double xi[D], const double xs[D], uint32_t si, uint32_t j, float c, uint32_t k[D], int32_t kj[D]) {
            float f[D];
#pragma unroll
            for (int a = 0; a < D; a++) {
                if constexpr (decltype(with_images)::value) f[a] = sep_f16(separation16<D>(P, xi[a], xs[a], a));
                else f[a] = sep_f16(xs[a] - xi[a]);
            }
            if constexpr (MULTI) {
                const uint32_t sj = sp[j], p = si * 4 + sj;
                return pair_kick2<D, true>(f, pt->rc2[p], pt->sigma2[p], pt->six_eps4[p], pt->r2_fast[p], c, pt->kick_c[sj], k, kj);
            } else {
                return pair_kick2<D, false>(f, rc2_0, sigma2_0, six_eps4_0, r2f_0, c, c, k, kj);
            }
        };
        auto own_of = [&](uint32_t own, uint32_t& si, float& c) {
            si = 0u;
            c = P.kick_c[0];
            if (MULTI) {
                si = sp[base + own];
                c = pt->kick_c[si];
            }
        };
        // In-block chunks: owner and partners both in LDS, every kick to the LDS accumulators.
        auto inblock_pass = [&](auto with_images) {
            const uint32_t nch = wfirst[0][NW];
            uint32_t g = t, w = 0u;
            auto slot = [&](uint32_t gg) -> size_t {
                while (gg >= wfirst[0][w + 1u]) w++;
                return (size_t)w * capw + (gg - wfirst[0][w]);
            };
            // the next chunk is in flight while this one is evaluated
            uint4 q_next = make_uint4(0u, 0u, 0u, 0u);
            if (g < nch) q_next = chunks[slot(g)];
            for (; g < nch; g += NT) {
                const uint4 qc = q_next;
                if (g + NT < nch) q_next = chunks[slot(g + NT)];
                uint32_t e[8], own, m;
                unpack_inblock(qc, e, own, m);
                double xi[D];
#pragma unroll
                for (int a = 0; a < D; a++) xi[a] = lpos[a][own];
                uint32_t si;
                float c;
                own_of(own, si, c);
                uint32_t k[D];
#pragma unroll
                for (int a = 0; a < D; a++) k[a] = 0u;
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++) {
                    if ((uint32_t)s2 >= m) break;
                    const uint32_t lj = e[s2];
                    double xs[D];
#pragma unroll
                    for (int a = 0; a < D; a++) xs[a] = lpos[a][lj];
                    int32_t kj[D];
                    if (pair(with_images, xi, xs, si, base + lj, c, k, kj)) {
#pragma unroll
                        for (int a = 0; a < D; a++) atomicAdd(&acc[a][lj], (uint32_t)kj[a]);
                    }
                }
#pragma unroll
                for (int a = 0; a < D; a++) atomicAdd(&acc[a][own], k[a]);
            }
        };
        // Cross-block chunks: partners gathered from the packed positions, owner kicks only (each
        // such pair is listed at both its ends).
        auto cross_pass = [&](auto with_images) {
            const uint32_t nch = wfirst[1][NW];
            uint32_t g = t, w = 0u;
            auto slot = [&](uint32_t gg) -> size_t {
                while (gg >= wfirst[1][w + 1u]) w++;
                return (size_t)w * capw + capw - 1u - (gg - wfirst[1][w]);
            };
            uint4 q_next = make_uint4(0u, 0u, 0u, 0u);
            if (g < nch) q_next = chunks[slot(g)];
            for (; g < nch; g += NT) {
                const uint4 qc = q_next;
                if (g + NT < nch) q_next = chunks[slot(g + NT)];
                uint32_t own, m;
                bool inblock;
                const uint4 q = untag_chunk(qc, own, m, inblock);
                const uint32_t i = base + own;
                double xi[D];
#pragma unroll
                for (int a = 0; a < D; a++) xi[a] = lpos[a][own];
                uint32_t si;
                float c;
                own_of(own, si, c);
                uint32_t k[D];
#pragma unroll
                for (int a = 0; a < D; a++) k[a] = 0u;
                const uint32_t js[4] = {q.x, m > 1u ? q.y : i, m > 2u ? q.z : i, m > 3u ? q.w : i};
                double xs[4][D];
#pragma unroll
                for (int e2 = 0; e2 < 4; e2++) {
                    if constexpr (D == 3) {
                        const double4 q4 = reinterpret_cast<const double4*>(B.aos[pp])[js[e2]];
                        xs[e2][0] = q4.x;
                        xs[e2][1] = q4.y;
                        xs[e2][2] = q4.z;
                    } else {
                        const double2 q2 = reinterpret_cast<const double2*>(B.aos[pp])[js[e2]];
                        xs[e2][0] = q2.x;
                        xs[e2][1] = q2.y;
                    }
                }
#pragma unroll
                for (int e2 = 0; e2 < 4; e2++) {
                    if ((uint32_t)e2 >= m) break;
                    // list separations stay inside the 8 sigma range (a system whose separations
                    // may reach it runs the full-list kernel, which tests the range)
                    int32_t kj[D];
                    (void)pair(with_images, xi, xs[e2], si, js[e2], c, k, kj);
                }
#pragma unroll
                for (int a = 0; a < D; a++) atomicAdd(&acc[a][own], k[a]);
            }
        };
        if (images) {
            inblock_pass(std::true_type{});
            cross_pass(std::true_type{});
        } else {
            inblock_pass(std::false_type{});
            cross_pass(std::false_type{});
        }
    }
    __syncthreads();
    for (uint32_t r = 0; r < PER; r++) {
        const uint32_t l = t + r * NT;
        const uint32_t i = base + l;
        if (i >= P.n) break;
        uint32_t si = 0;
        float c = P.kick_c[0];
        if (MULTI) {
            si = sp[i];
            c = pt->kick_c[si];
        }
        uint32_t k[D];
        double xi[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            k[a] = 0u;
            xi[a] = lpos[a][l] * 16.0;
        }
        float wf[2 * D] = {};
        if (P.soft_walls) soft_walls<D>(P, xi, c, k, nullptr, wf);
#pragma unroll
        for (int a = 0; a < D; a++) k[a] += acc[a][l];
        int32_t v[D];
        double x[D];
        advance<D, BOOK>(P, B, i, pm, si, k, xi, ctl[CTL_CLOCK], wf, x, v);
        float r2 = 0.0f;
        bool far = false;
#pragma unroll
        for (int a = 0; a < D; a++) {
            B.pos[pp ^ 1][a][i] = x[a];
            B.vel[pm][a][i] = v[a];
            float d;
            far |= !sep_f(separation<D>(P, B.ref[a][i], x[a], a), d);
            r2 += d * d;
        }
        if constexpr (D == 3) reinterpret_cast<double4*>(B.aos[pp ^ 1])[i] = make_double4(x[0] * 0.0625, x[1] * 0.0625, x[2] * 0.0625, 0.0);
        else reinterpret_cast<double2*>(B.aos[pp ^ 1])[i] = make_double2(x[0] * 0.0625, x[1] * 0.0625);
        bool stale = far || r2 >= P.half_skin2;
        uint64_t bal = ballot64(stale);
        if (bal != 0ull && __lane_id() == (uint32_t)(__ffsll((long long)bal) - 1) && *(volatile const uint32_t*)&B.ctl[CTL_STALE_NEXT] == 0u)
            atomicOr(&B.ctl[CTL_STALE_NEXT], 1u);
    }
}

// ---------------------------------------------------------------------------------------------
// Cell-list forces (ForceMode::CellList, and kicks()/reverse()/energy readouts)
// ---------------------------------------------------------------------------------------------

// md_core::table::IntTables::kick (table.rs:75-105) on an exact separation d = r_j - r_i (2^-32
// sigma units, held in f64): magnitudes rounded to 2^-20 sigma, r2 to 2^-24 sigma^2, the table
// interpolated linearly in integers, kick = -T q >> (T_BITS + SEP_BITS) on magnitudes, then signed.
template <int D>
__device__ __forceinline__ bool int_kick2(const Params<D>& P, const Bufs<D>& B, uint32_t sa, uint32_t sb, double dx, double dy, int32_t out[2]) {
    const double limit = 34359738368.0;  // 8 sigma
    if (!(dx >= -limit && dx < limit && dy >= -limit && dy < limit)) return false;
    // (|d| + 2^11) >> 12 = floor(|d| / 2^12 + 1/2): exact in f64 for |d| < 2^35
    const int64_t mx = (int64_t)__builtin_floor(__builtin_fabs(dx) * 0x1p-12 + 0.5);
    const int64_t my = (int64_t)__builtin_floor(__builtin_fabs(dy) * 0x1p-12 + 0.5);
    const int64_t q[2] = {dx < 0.0 ? -mx : mx, dy < 0.0 ? -my : my};
    const uint32_t r2 = (uint32_t)(((uint64_t)(q[0] * q[0] + q[1] * q[1]) + (uint64_t(1) << 15)) >> 16);
    const uint32_t pair = sa * P.nspecies + sb;
    if (r2 >= B.irc2[pair]) return false;
    const uint32_t idx = r2 >> 12;
    const int64_t frac = (int64_t)(r2 & 4095u);
    const int32_t* tab = B.itab + (size_t)pair * P.itab_len;
    const int64_t t0 = tab[idx], t1 = tab[idx + 1];
    const int64_t t = t0 + (((t1 - t0) * frac) >> 12);
    const uint64_t at = t < 0 ? (uint64_t)0 - (uint64_t)t : (uint64_t)t;
#pragma unroll
    for (int a = 0; a < 2; a++) {
        const uint64_t aq = q[a] < 0 ? (uint64_t)0 - (uint64_t)q[a] : (uint64_t)q[a];
        const int64_t m = (int64_t)((at * aq) >> 22);  // |t| < 2^31, |q| < 2^23: fits 64 bits
        const bool neg = (t < 0) != (q[a] < 0);
        out[a] = (int32_t)(uint32_t)(uint64_t)(neg ? m : -m);
    }
    return true;
}

// Pair::energy in f64 (pair.rs:36-43), for the energy readout of the integer forces.
__device__ __forceinline__ double pair_energy_f64(const PairTable* pt, uint32_t p, double r2) {
    if (r2 >= pt->rc2_64[p]) return 0.0;
    const double s2 = pt->sigma2_64[p] / r2;
    const double s6 = s2 * s2 * s2;
    return pt->eps4_64[p] * s6 * (s6 - 1.0) - pt->shift_64[p];
}

// items: slot of each sorted position (null: slots are already sorted, items[q] = q).
template <int D, bool MULTI>
__global__ __launch_bounds__(TPB) void k_force_cells(Params<D> P, Bufs<D> B, const uint32_t* items, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS], pm = B.ctl[CTL_PMETA];
    const double* px[D];
    for (int a = 0; a < D; a++) px[a] = B.pos[pp][a];
    const uint32_t* sp = B.species[pm];
    double xi[D];
    for (int a = 0; a < D; a++) xi[a] = px[a][i];
    uint32_t si = MULTI ? sp[i] : 0u;
    float c = B.pt->kick_c[si];
    uint32_t k[D];
    for (int a = 0; a < D; a++) k[a] = 0u;
    double pe = 0.0, wall_pe = 0.0;
    uint32_t cell[D];
    cell_of<D>(P, xi, cell);
    for (int o = 0; o < stencil_size(D); o++) {
        if (stencil_dup<D>(P, o)) continue;
        uint32_t cc[D];
        stencil_cell<D>(P, cell, o, cc);
        uint32_t key = morton<D>(cc);
        uint32_t s = B.cell_start[key], e = B.cell_end[key];
        for (uint32_t q = s; q < e; q++) {
            uint32_t j = items ? items[q] : q;
            if (j == i) continue;
            if constexpr (D == 2) {
                // System::interact_across (system.rs:229-231): insulating compartment walls
                if (P.comp_filter && (__builtin_floor(px[0][j] * P.comp_scale[0]) != __builtin_floor(xi[0] * P.comp_scale[0]) ||
                                      __builtin_floor(px[1][j] * P.comp_scale[1]) != __builtin_floor(xi[1] * P.comp_scale[1])))
                    continue;
                if (P.arith) {
                    const uint32_t sj = MULTI ? sp[j] : 0u;
                    const double dx = separation<D>(P, xi[0], px[0][j], 0), dy = separation<D>(P, xi[1], px[1][j], 1);
                    int32_t kk[2];
                    if (int_kick2<D>(P, B, si, sj, dx, dy, kk)) {
                        k[0] += (uint32_t)kk[0];
                        k[1] += (uint32_t)kk[1];
                        const double fx = dx * 0x1p-32, fy = dy * 0x1p-32;
                        pe += pair_energy_f64(B.pt, si * 4 + sj, fx * fx + fy * fy);
                    }
                    continue;
                }
            }
            float f[D];
            bool ok = true;
#pragma unroll
            for (int a = 0; a < D; a++) ok &= sep_f(separation<D>(P, xi[a], px[a][j], a), f[a]);
            uint32_t p = MULTI ? si * 4 + sp[j] : 0u;
            if (ok) pair_kick_pe<D>(f, B.pt, (int)p, c, k, pe);
        }
    }
    if (P.soft_walls && !P.feat) soft_walls<D>(P, xi, c, k, &wall_pe);  // the boundary path adds them after
    for (int a = 0; a < D; a++) B.kicks[a][i] = (int32_t)k[a];
    B.pe[i] = 0.5 * pe + wall_pe;
}

// Leapfrog in place from the stored kicks (cell-list mode).
// `clock_add`: steps already taken in this step() call (the clock word advances at its end).
template <int D>
__global__ __launch_bounds__(TPB) void k_integrate(Params<D> P, Bufs<D> B, uint32_t clock_add) {
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS], pm = B.ctl[CTL_PMETA];
    uint32_t si = B.species[pm][i];
    double xi[D];
    uint32_t k[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        xi[a] = B.pos[pp][a][i];
        k[a] = (uint32_t)B.kicks[a][i];
    }
    float wf[2 * D] = {};
    if (P.ledger && P.soft_walls) {
        // the soft-wall forces of this step's kicks, for the ledger (the kicks already hold them)
        uint32_t dummy[D] = {};
        soft_walls<D>(P, xi, B.pt->kick_c[si], dummy, nullptr, wf);
    }
    int32_t v[D];
    double x[D];
    advance<D, true>(P, B, i, pm, si, k, xi, B.ctl[CTL_CLOCK] + clock_add, wf, x, v);
#pragma unroll
    for (int a = 0; a < D; a++) {
        B.pos[pp][a][i] = x[a];
        B.vel[pm][a][i] = v[a];
    }
}

// v <- -(v + kick + g) (int_engine3.rs:209-219)
template <int D>
__global__ __launch_bounds__(TPB) void k_reverse(Params<D> P, Bufs<D> B) {
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pm = B.ctl[CTL_PMETA];
    uint32_t si = B.species[pm][i];
    const bool pin = (P.pinned >> si) & 1u;
#pragma unroll
    for (int a = 0; a < D; a++) {
        uint32_t v = (uint32_t)B.vel[pm][a][i] + (uint32_t)B.kicks[a][i] + (pin ? 0u : (uint32_t)P.gravity[a]);
        B.vel[pm][a][i] = (int32_t)((uint32_t)0 - v);
    }
}

// Half-step kinetic energy per slot: 0.5 m sum (v UNIT / dt)^2, zero when pinned.
template <int D>
__global__ __launch_bounds__(TPB) void k_kinetic(Params<D> P, Bufs<D> B, double m0, double m1, double m2, double m3, double dt) {
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pm = B.ctl[CTL_PMETA];
    uint32_t si = B.species[pm][i];
    double m = si == 0 ? m0 : si == 1 ? m1 : si == 2 ? m2 : m3;
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        double u = (double)B.vel[pm][a][i] * (1.0 / 4294967296.0) / dt;
        s += u * u;
    }
    B.ke[i] = ((P.pinned >> si) & 1u) ? 0.0 : 0.5 * m * s;
}

// md3.wgsl kick_region (a GPU-only command; f32 coordinates and round-half-even as there).
template <int D>
__global__ __launch_bounds__(TPB) void k_kick_region(Params<D> P, Bufs<D> B, float lo0, float lo1, float lo2, float hi0, float hi1, float hi2, float scale, int32_t a0, int32_t a1, int32_t a2) {
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS], pm = B.ctl[CTL_PMETA];
    if ((P.pinned >> B.species[pm][i]) & 1u) return;
    float lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    int32_t add[3] = {a0, a1, a2};
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double x = B.pos[pp][a][i];
        const double whole = __builtin_floor(x * 0x1p-32);  // the i64's two 32-bit words, exactly
        float xf = (float)(uint32_t)whole + __builtin_ldexpf((float)(uint32_t)(x - whole * 4294967296.0), -32);
        if (xf < lo[a] || xf >= hi[a]) return;
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        float v = (float)B.vel[pm][a][i] * scale;
        B.vel[pm][a][i] = (int32_t)((uint32_t)cvt_sat_i32(__builtin_rintf(v)) + (uint32_t)add[a]);
    }
}

template <int D>
__global__ __launch_bounds__(TPB) void k_set_species(Params<D> P, Bufs<D> B, const uint32_t* list, uint32_t count, uint32_t to) {
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n || count == 0) return;
    const uint32_t pm = B.ctl[CTL_PMETA];
    uint32_t id = B.ids[pm][i];
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (list[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    if (lo < count && list[lo] == id) B.species[pm][i] = to;
}

template <int D>
__global__ __launch_bounds__(TPB) void k_set_species_velocity(Params<D> P, Bufs<D> B, uint32_t s, int32_t a0, int32_t a1, int32_t a2) {
    uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pm = B.ctl[CTL_PMETA];
    if (B.species[pm][i] != s) return;
    int32_t add[3] = {a0, a1, a2};
#pragma unroll
    for (int a = 0; a < D; a++) B.vel[pm][a][i] = add[a];
}

// ---------------------------------------------------------------------------------------------
// The 2D boundary path (md_core::System's compartments, membrane, thermostat regions, heat-bath
// walls, piston and bonds; IntEngine::forces / step, int_engine.rs:130-361).  One step is four
// launches: pair kicks (cell list or full Verlet list), the per-particle terms (bonds, soft
// walls, piston), the leapfrog with every boundary, and the piston's own move.
// ---------------------------------------------------------------------------------------------

constexpr int LEDGER2_COMPARTMENTS = 256;  // md_core::system::MAX_COMPARTMENTS
constexpr double PISTON_FORCE_SCALE = 65536.0;  // walls::PISTON_FORCE_SCALE

// Compartment coordinate of a position along an axis (System::compartment_of, system.rs:244-247):
// x >> shift, computed as floor(x 2^-shift), exact for the f64 integer positions.
__device__ __forceinline__ double compartment_of2(const Params<2>& P, double x, int a) { return __builtin_floor(x * P.comp_scale[a]); }

// Book a wall exchange in the totals and in the entry of compartment c along the wall (only
// compartments 0..255 have entries, as in CompartmentLedger; a position outside the box, as a
// piston below zero can make, books in the totals alone).
__device__ __forceinline__ void book2(const Bufs<2>& B, int w, double c, int field, double value) {
    ledger_add(B.ledger, w, field, value);
    if (c >= 0.0 && c < (double)LEDGER2_COMPARTMENTS) ledger_add(B.cledger + 3 * (w * LEDGER2_COMPARTMENTS + (int)c), 0, field, value);
}

__device__ __forceinline__ void book2_hit(const Bufs<2>& B, int w, double c) {
    atomicAdd(&B.ledger[3 * w + 2], 1ull);
    if (c >= 0.0 && c < (double)LEDGER2_COMPARTMENTS) atomicAdd(&B.cledger[3 * (w * LEDGER2_COMPARTMENTS + (int)c) + 2], 1ull);
}

// fixed::sep_long_f32 (fixed.rs:65-72): the short-range conversion below 8 sigma, else the high
// and low words of |d| converted separately and added in f32 (bonds may stretch beyond 8 sigma).
__device__ __forceinline__ float sep_long_f(double d) {
    float f;
    if (sep_f(d, f)) return f;
    const uint64_t m = (uint64_t)__builtin_fabs(d);
    const float mag = (float)(uint32_t)(m >> 32) + (float)(uint32_t)m * (1.0f / 4294967296.0f);
    return d < 0.0 ? -mag : mag;
}

// Pair kicks over the full Verlet list, without the leapfrog (the boundary path's list mode).
template <bool MULTI>
__global__ __launch_bounds__(TPB) void k_force_list2(Params<2> P, Bufs<2> B) {
    const uint32_t* ctl = B.ctl;
    if (!step_runs(ctl)) return;
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = ctl[CTL_PPOS], pm = ctl[CTL_PMETA];
    const double* px = B.pos[pp][0];
    const double* py = B.pos[pp][1];
    const uint32_t* sp = B.species[pm];
    const double xi[2] = {px[i], py[i]};
    const uint32_t si = MULTI ? sp[i] : 0u;
    const PairTable* __restrict__ pt = B.pt;
    const float c = pt->kick_c[si];
    const double ci0 = compartment_of2(P, xi[0], 0), ci1 = compartment_of2(P, xi[1], 1);
    uint32_t k[2] = {0u, 0u};
    // A valid full list never exceeds cap; a count beyond it (a list left in another form) is
    // reported as an overflow, which stops the steps until the list is regrown and rebuilt,
    // and only the cap entries that exist are read.
    const uint32_t cnt = B.nbr_count[i];
    if (cnt > P.cap) atomicOr(&B.ctl[CTL_OVERFLOW], 1u);
    const uint32_t m = min(cnt, P.cap);
    const uint32_t* nb = B.nbr + i;
    for (uint32_t e = 0; e < m; e++) {
        const uint32_t j = nb[(size_t)e * P.n];
        const double xj[2] = {px[j], py[j]};
        // System::interact_across (system.rs:229-231)
        if (P.comp_filter && (compartment_of2(P, xj[0], 0) != ci0 || compartment_of2(P, xj[1], 1) != ci1)) continue;
        const double dx = separation<2>(P, xi[0], xj[0], 0), dy = separation<2>(P, xi[1], xj[1], 1);
        if (P.arith) {
            // md::Arith::Int: the all-integer table kick
            int32_t kk[2];
            if (int_kick2<2>(P, B, si, MULTI ? sp[j] : 0u, dx, dy, kk)) {
                k[0] += (uint32_t)kk[0];
                k[1] += (uint32_t)kk[1];
            }
            continue;
        }
        float f[2];
        bool ok = sep_f(dx, f[0]);
        ok &= sep_f(dy, f[1]);
        const uint32_t p = MULTI ? si * 4 + sp[j] : 0u;
        if (ok) pair_kick<2>(f, pt->rc2[p], pt->sigma2[p], pt->six_eps4[p], pt->r2_fast[p], c, k);
    }
    B.kicks[0][i] = (int32_t)k[0];
    B.kicks[1][i] = (int32_t)k[1];
}

enum : uint32_t { EX_GATED = 1, EX_PE = 2, EX_PISTON = 4, EX_BOOK = 8 };

// The per-particle terms after the pairs (int_engine.rs:181-248): bonds (BondType::eval in f32
// on sep_long separations, rounded per bond), soft walls (their force booked as impulse x dt in
// the ledger of the compartment along the wall), and the piston, a soft wall at its height
// pushing down (its gas force summed in 2^-16 units with i64 atomics, so in any order).
__global__ __launch_bounds__(TPB) void k_extras2(Params<2> P, Bufs<2> B, uint32_t flags) {
    const uint32_t* ctl = B.ctl;
    if ((flags & EX_GATED) && !step_runs(ctl)) return;
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = ctl[CTL_PPOS], pm = ctl[CTL_PMETA];
    const double xi[2] = {B.pos[pp][0][i], B.pos[pp][1][i]};
    const uint32_t si = B.species[pm][i];
    const float c = B.pt->kick_c[si];
    uint32_t k[2] = {(uint32_t)B.kicks[0][i], (uint32_t)B.kicks[1][i]};
    double pe = 0.0;
    if (P.bonds) {
        const uint32_t* row = B.partners + 8 * (size_t)B.ids[pm][i];
        float bond_pe = 0.0f;
        for (int b = 0; b < 4; b++) {
            const uint32_t jid = row[2 * b];
            if (jid == UINT32_MAX) break;
            const uint32_t t = row[2 * b + 1];
            const uint32_t j = B.slot_of[jid];
            const float fx = sep_long_f(separation<2>(P, xi[0], B.pos[pp][0][j], 0));
            const float fy = sep_long_f(separation<2>(P, xi[1], B.pos[pp][1][j], 1));
            const float r2 = fx * fx + fy * fy;
            const float bk = P.bond_k[t], br0 = P.bond_r0[t];
            float fr, u;
            if ((P.bond_fene >> t) & 1u) {
                float x = r2 / (br0 * br0);
                if (x > 0.999999f) x = 0.999999f;
                u = -0.5f * bk * br0 * br0 * (float)log((double)(1.0f - x));
                fr = -bk / (1.0f - x);
            } else if (br0 == 0.0f) {
                fr = -bk;
                u = 0.5f * bk * r2;
            } else {
                const float r = __builtin_sqrtf(r2);
                const float dr = r - br0;
                fr = -bk * dr / r;
                u = 0.5f * bk * dr * dr;
            }
            const float a = -fr * c;
            k[0] += (uint32_t)cvt_sat_i32(__builtin_rintf(a * fx));
            k[1] += (uint32_t)cvt_sat_i32(__builtin_rintf(a * fy));
            bond_pe += u;
        }
        pe += 0.5 * (double)bond_pe;
    }
    if (P.soft_walls) {
        float wf[4] = {0.f, 0.f, 0.f, 0.f};
        soft_walls<2>(P, xi, c, k, &pe, wf);
        if (flags & EX_BOOK) {
#pragma unroll
            for (int w = 0; w < 4; w++)
                if (wf[w] != 0.0f) {
                    const int a = w / 2;
                    book2(B, w, compartment_of2(P, xi[1 - a], 1 - a), 1, (double)wf[w] * P.dt);
                }
        }
    }
    if (P.piston) {
        float df;
        if (sep_f((double)B.piston->w - xi[1], df) && df < SOFT_RC_F) {
            const float s2 = 1.0f / (df * df);
            const float s6 = s2 * s2 * s2;
            const float f = 24.0f * s6 * (2.0f * s6 - 1.0f) / df;
            k[1] -= (uint32_t)cvt_sat_i32(__builtin_rintf(f * c));
            if (flags & EX_PISTON) atomicAdd(&B.piston->force_sum, (unsigned long long)(long long)round((double)f * PISTON_FORCE_SCALE));
            pe += (double)(4.0f * s6 * (s6 - 1.0f) + 1.0f);
        }
    }
    B.kicks[0][i] = (int32_t)k[0];
    B.kicks[1][i] = (int32_t)k[1];
    if (flags & EX_PE) B.pe[i] += pe;
}

// walls::piston_kick (walls.rs:143-146): a displacement-per-step change from a net force.
__device__ __forceinline__ int32_t piston_kick2(const Params<2>& P, double net) {
    const double x = round(net * P.dt * P.dt / P.piston_mass * 4294967296.0);
    if (x != x) return 0;
    return x >= 2147483647.0 ? INT32_MAX : (x <= -2147483648.0 ? INT32_MIN : (int32_t)x);
}

// The leapfrog with md_core::System's boundaries (int_engine.rs:266-337): v += kick + g; the
// thermostat in its region (walls::andersen, walls.rs:120-139, with the fan's drift and the
// affine flow under a driven piston); x += v; the membrane (walls.rs:17-24, pores
// system.rs:116-119); walls (hard, soft, heat bath, walls.rs:27-35, 91-104) or the compartment
// walls (walls.rs:76-86); the piston's mirror.  Exchanges go to the ledger under the compartment
// along the wall, taken at the start of the step.  LIST: gated on the list, writes the other
// position buffer and checks the skin; else in place.
template <bool LIST>
__global__ __launch_bounds__(TPB) void k_advance2(Params<2> P, Bufs<2> B, uint32_t clock_add) {
    const uint32_t* ctl = B.ctl;
    if (LIST && !step_runs(ctl)) return;
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = ctl[CTL_PPOS], pm = ctl[CTL_PMETA];
    const uint32_t si = B.species[pm][i];
    const uint32_t id = B.ids[pm][i];
    const uint32_t clock = ctl[CTL_CLOCK] + clock_add;
    const bool pin = (P.pinned >> si) & 1u;
    const double m = P.mass[si];
    const double xi[2] = {B.pos[pp][0][i], B.pos[pp][1][i]};
    int32_t v[2];
#pragma unroll
    for (int a = 0; a < 2; a++) v[a] = (int32_t)((uint32_t)B.vel[pm][a][i] + (uint32_t)B.kicks[a][i] + (pin ? 0u : (uint32_t)P.gravity[a]));
    const double home[2] = {compartment_of2(P, xi[0], 0), compartment_of2(P, xi[1], 1)};
    double pw = 0.0;
    int32_t pv = 0;
    if (P.piston) {
        pw = (double)B.piston->w;
        pv = B.piston->v;
    }
    const double xb = P.bath_axis ? xi[1] : xi[0];
    if (P.bath && !pin && ((P.bath_species >> si) & 1u) && xb >= P.bath_lo && xb < P.bath_hi) {
        double u[2];
        if (thermostat_draw2(P, id, clock, si, u)) {
            const double b0 = velocity(P, v[0]), b1 = velocity(P, v[1]);
            v[0] = to_step(P, u[0]);
            v[1] = to_step(P, u[1]);
            v[0] = (int32_t)((uint32_t)v[0] + (uint32_t)P.bath_drift);
            if (P.piston && !P.piston_free) {
                const double f = fmin(fmax((xi[1] * 0x1p-32) / (pw * 0x1p-32), 0.0), 1.0);
                const double q = round((double)pv * f);
                const int32_t add = q != q ? 0 : (q >= 2147483647.0 ? INT32_MAX : (q <= -2147483648.0 ? INT32_MIN : (int32_t)q));
                v[1] = (int32_t)((uint32_t)v[1] + (uint32_t)add);
            }
            const double a0 = velocity(P, v[0]), a1 = velocity(P, v[1]);
            book2(B, 6, 0.0, 0, 0.5 * m * (a0 * a0 + a1 * a1 - b0 * b0 - b1 * b1));
            book2_hit(B, 6, 0.0);
        }
    }
    double x[2] = {xi[0] + (double)v[0], xi[1] + (double)v[1]};
    if (P.membrane && ((P.mem_blocks >> si) & 1u)) {
        // across axis a (selects instead of indexing, so the arrays stay in registers)
        const bool ay = P.mem_axis != 0u;
        const double old_a = ay ? xi[1] : xi[0], new_a = ay ? x[1] : x[0], along = ay ? xi[0] : xi[1];
        const bool in_pore = P.pore_shift != 0u && ((long long)along & P.pore_mask) < P.pore_width;
        const bool lo_old = old_a < P.mem_pos, lo_new = new_a < P.mem_pos;
        if (!in_pore && lo_old != lo_new) {
            const int face = lo_old ? 4 : 5;
            const double mirrored = 2.0 * P.mem_pos - 1.0 - new_a;
            const int32_t va = ay ? v[1] : v[0];
            const double dp = 2.0 * m * __builtin_fabs(velocity(P, va));
            const int32_t flipped = (int32_t)((uint32_t)0 - (uint32_t)va);
            if (ay) {
                x[1] = mirrored;
                v[1] = flipped;
            } else {
                x[0] = mirrored;
                v[0] = flipped;
            }
            const double cc = compartment_of2(P, along, ay ? 0 : 1);
            book2(B, face, cc, 1, dp);
            book2_hit(B, face, cc);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; a++) {
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
        if (side >= 0) {
            const int w = 2 * a + side;
            const double b0 = velocity(P, v[0]), b1 = velocity(P, v[1]);
            if ((P.thermal_walls >> w) & 1u) {
                double vn, t;
                thermal_wall_draw2(P, id, clock, w, side, si, vn, t);
                v[a] = to_step(P, vn);
                v[1 - a] = to_step(P, t);
            } else {
                v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
            }
            const double a0 = velocity(P, v[0]), a1 = velocity(P, v[1]);
            const double ka = 0.5 * m * (a0 * a0 + a1 * a1), kb = 0.5 * m * (b0 * b0 + b1 * b1);
            const double before = a == 0 ? b0 : b1, after = a == 0 ? a0 : a1;
            const double cc = compartment_of2(P, xi[1 - a], 1 - a);
            book2(B, w, cc, 0, ka - kb);
            book2(B, w, cc, 1, __builtin_fabs(m * (before - after)));
            book2_hit(B, w, cc);
        } else if (P.comp_shift[a] != 0u && compartment_of2(P, x[a], a) != home[a]) {
            // walls::mirror_compartment: reflect off the wall of the home compartment
            const double side_len = 1.0 / P.comp_scale[a];
            const double lo = home[a] * side_len, hi = (home[a] + 1.0) * side_len;
            x[a] = x[a] < lo ? 2.0 * lo - 1.0 - x[a] : (x[a] >= hi ? 2.0 * hi - 1.0 - x[a] : x[a]);
            v[a] = (int32_t)((uint32_t)0 - (uint32_t)v[a]);
        }
    }
    if (P.piston && x[1] >= pw) {
        x[1] = 2.0 * pw - 1.0 - x[1];
        v[1] = (int32_t)((uint32_t)0 - (uint32_t)v[1]);
    }
    const uint32_t out = LIST ? pp ^ 1u : pp;
#pragma unroll
    for (int a = 0; a < 2; a++) {
        B.pos[out][a][i] = x[a];
        B.vel[pm][a][i] = v[a];
    }
    if (LIST) {
        float r2 = 0.0f;
        bool far = false;
#pragma unroll
        for (int a = 0; a < 2; a++) {
            float d;
            far |= !sep_f(separation<2>(P, B.ref[a][i], x[a], a), d);
            r2 += d * d;
        }
        const bool stale = far || r2 >= P.half_skin2;
        const uint64_t bal = ballot64(stale);
        if (bal != 0ull && __lane_id() == (uint32_t)(__ffsll((long long)bal) - 1) && *(volatile const uint32_t*)&B.ctl[CTL_STALE_NEXT] == 0u)
        atomicOr(&B.ctl[CTL_STALE_NEXT], 1u);
    }
}

// IntEngine::move_piston (int_engine.rs:346-361): a free piston's kick from the net force, the
// drift, the stops, and the work the gas did on it.  One lane.
__global__ void k_piston2(Params<2> P, Bufs<2> B, int gated) {
    if (gated && !step_runs(B.ctl)) return;
    PistonDev& p = *B.piston;
    const double force = (double)(long long)p.force_sum / PISTON_FORCE_SCALE;
    p.force_sum = 0ull;
    if (P.piston_free) p.v = (int32_t)((uint32_t)p.v + (uint32_t)piston_kick2(P, force - P.piston_load));
    const long long old = p.w;
    p.w += (long long)p.v;
    if (p.w < P.piston_lo || p.w > P.piston_hi) {
        p.w = p.w < P.piston_lo ? P.piston_lo : P.piston_hi;
        p.v = 0;
    }
    p.work += force * ((double)(p.w - old) * 0x1p-32);
    p.force = force;
}

// IntEngine::reverse's piston (int_engine.rs:405-411): v <- -(v + kick).  One lane.
__global__ void k_piston_reverse2(Params<2> P, Bufs<2> B) {
    PistonDev& p = *B.piston;
    const double force = (double)(long long)p.force_sum / PISTON_FORCE_SCALE;
    p.force_sum = 0ull;
    const int32_t kick = P.piston_free ? piston_kick2(P, force - P.piston_load) : 0;
    p.v = (int32_t)((uint32_t)0 - ((uint32_t)p.v + (uint32_t)kick));
}

// The id -> slot map of the freshly gathered order (before the rebuild commits it).
__global__ __launch_bounds__(TPB) void k_slot_of(Params<2> P, Bufs<2> B, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    const uint32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= P.n) return;
    B.slot_of[B.ids[B.ctl[CTL_PMETA] ^ 1u][k]] = k;
}

// GpuEngine::flip_bit / flip_species_bit (engine.rs:964-973, 1014-1023): XOR bit `bit` of
// coordinate `axis` of the particle with id `sel` (by_species 0) or of every particle whose
// species is in the bitmask `sel`.  check = 1 only counts flips that would leave the box.
__global__ __launch_bounds__(TPB) void k_flip2(Params<2> P, Bufs<2> B, uint32_t sel, int by_species, uint32_t axis, uint32_t bit, int check) {
    const uint32_t i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t pp = B.ctl[CTL_PPOS], pm = B.ctl[CTL_PMETA];
    const bool hit = by_species ? ((sel >> B.species[pm][i]) & 1u) != 0u : B.ids[pm][i] == sel;
    if (!hit) return;
    const long long x = (long long)B.pos[pp][axis][i] ^ (1ll << bit);
    if (check) {
        if (x < 0 || x >= P.box[axis]) atomicAdd(&B.ctl[CTL_FLIP_BAD], 1u);
        return;
    }
    B.pos[pp][axis][i] = (double)x;
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------

// A 2D system with any of md_core::System's boundaries beyond hard and soft walls: it runs on
// the split path.
static bool boundary_features2(const SystemDesc& sys) {
    bool thermal = false;
    for (int a = 0; a < 2; a++)
        for (int side = 0; side < 2; side++) thermal |= sys.walled[a] && sys.wall_kind[a][side] == RH_WALL_THERMAL;
    return sys.compartments[0] || sys.compartments[1] || sys.has_membrane || sys.has_piston || !sys.bonds.empty() || sys.nbond_types > 0 ||
           sys.has_bath || thermal;
}

static bool has_soft_walls(const SystemDesc& sys) {
    bool soft = false;
    for (int a = 0; a < sys.dim; a++)
        for (int side = 0; side < 2; side++) soft |= sys.walled[a] && sys.wall_kind[a][side] == RH_WALL_SOFT;
    return soft;
}

template <int D>
void* Engine<D>::dalloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) throw Error(std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    allocs_.push_back(p);
    alloc_bytes_ += bytes ? bytes : 16;
    return p;
}

// Plans an engine without allocating anything: the checks and parameters of the constructor.
template <int D>
Engine<D>::Engine(const SystemDesc& sys, uint64_t n, const rh_mode& mode) : sys_(sys), mode_(mode), n_(n) {
    RH_REQUIRE(n >= 1, "an engine needs at least one particle");
    RH_REQUIRE(n < (uint64_t(1) << 31), "too many particles for one engine");
    RH_REQUIRE(mode.kind <= 1, "unknown force mode");
    if (mode.kind == 1) RH_REQUIRE(mode.skin >= 0.0 && mode.cap >= 1, "Verlet mode needs skin >= 0 and cap >= 1");
    std::memset(&B_, 0, sizeof B_);
    block_ = n >= (uint64_t(1) << 20) ? 2048u : (n >= (uint64_t(1) << 18) ? 1024u : 512u);
    build_params(sys, mode);
}

// The device memory the constructor allocates for this plan (the sum of its allocations below,
// with the list at its initial capacity); an engine's own total is alloc_bytes().
template <int D>
uint64_t Engine<D>::planned_bytes() const {
    const uint64_t n = n_;
    auto sz = [](uint64_t b) { return b ? b : uint64_t(16); };
    uint64_t t = 0;
    t += 2 * (D * (sz(n * 8) + sz(n * 4)) + sz(n * 4) + sz(n * (D == 3 ? 4 : 2) * 8) + sz(n * 4));  // pos, vel, species, aos, ids
    t += D * (sz(n * 8) + sz(n * 4));                                                               // ref, kicks
    t += sz(n * 4);                                                                                  // nbr_count
    t += sz(((n + 4095u) & ~uint64_t(4095)) / 64u * 8);                                              // wave_chunks
    t += sz((n + 32u) * 16);                                                                         // rel
    if (P_.rows) t += sz((2 * n + 32u) * 16) + sz(row_bins_ * 4) + sz((row_keys_ + 1u) * 4) + sz(n * 4) + sz(scan_temp_bytes(row_keys_ + 1u)) + sz(scan_temp_bytes(ncell_keys_));
    t += 2 * sz(ncell_keys_ * 4);                                                                    // cell_start, cell_end
    t += 2 * sz(n * 8);                                                                              // pe, ke
    t += sz(CTL_WORDS * 4) + sz(LEDGER_WORDS * 8);
    if constexpr (D == 2) {
        if (P_.bonds) t += sz(n * 4) + sz(n * 32);
        if (P_.piston) t += sz(sizeof(PistonDev));
        if (P_.feat) t += sz(7 * LEDGER2_COMPARTMENTS * 3 * 8);
    }
    t += sz(reduce_temp_bytes(n)) + sz(64);
    if (mode_.kind == 1) {
        const uint64_t cap = (mode_.cap + 3u) & ~3u;
        t += sz(cap * ((n + 4095u) & ~uint64_t(4095)) * 4);
    }
    const uint64_t nb = blocks_for(n, 4096);  // the radix sort: two key/value pairs, histograms, scan
    t += 4 * sz(n * 4) + sz(256 * nb * 4) + sz(scan_temp_bytes(256 * nb));
    t += sz(sizeof(PairTable));
    return t;
}

template <int D>
void Engine<D>::build_params(const SystemDesc& sys, const rh_mode& mode) {
    RH_REQUIRE(sys.nspecies >= 1 && sys.nspecies <= 4, "between 1 and 4 species");
    Params<D>& P = P_;
    std::memset(&P, 0, sizeof P);
    P.n = (uint32_t)n_;
    P.nspecies = sys.nspecies;
    P.pinned = sys.pinned;
    double skin = mode.kind == 1 ? mode.skin : 0.0;
    double rcm = core::rc_max(sys.pairs, sys.nspecies);
    RH_REQUIRE(rcm < 7.9, "cutoff must be below 8 sigma");
    // Cells at least (rc + skin) wide plus a rounding margin (three.rs:151-157).
    double b[3] = {0, 0, 0}, bmax = 0.0;
    for (int a = 0; a < D; a++) {
        b[a] = core::to_f64(sys.box_len[a]);
        bmax = std::max(bmax, b[a]);
    }
    double margin = 1e-3 + 1e-6 * bmax;
    uint32_t ncmax = 1;
    for (int a = 0; a < D; a++) {
        P.nc[a] = core::sat_u32(std::floor(b[a] / (rcm + skin + margin)));
        if (P.nc[a] < 3) {
            // Fewer than three cells fit: one cell spans the axis (every pair along it is a
            // candidate).  The minimum image stays unique while L exceeds twice the list radius.
            RH_REQUIRE(b[a] > 2.0 * (rcm + skin) * (1.0 + 1e-4) + margin, "box too small: shorter than twice the cutoff plus skin");
            P.nc[a] = 1;
        }
        ncmax = std::max(ncmax, P.nc[a]);
        RH_REQUIRE(sys.box_len[a] < (int64_t(1) << 52), "box longer than 2^20 sigma");
        P.box[a] = sys.box_len[a];
        P.half[a] = sys.box_len[a] / 2;
        P.boxf[a] = (double)sys.box_len[a];
        P.inv_box[a] = sys.walled[a] ? 0.0 : 1.0 / P.boxf[a];  // 0 on walled axes: no image
        P.box16[a] = P.boxf[a] * 0.0625;
        P.inv_box16[a] = P.inv_box[a] * 16.0;
        P.cell_scale[a] = (double)P.nc[a] / (double)sys.box_len[a];
        if (!sys.walled[a]) P.periodic |= 1u << a;
        else
            for (int side = 0; side < 2; side++) {
                const int w = 2 * a + side;
                if (sys.wall_kind[a][side] == RH_WALL_SOFT) P.soft_walls |= 1u << w;
                if (sys.wall_kind[a][side] == RH_WALL_THERMAL) {
                    P.thermal_walls |= 1u << w;
                    for (uint32_t sp = 0; sp < sys.nspecies; sp++) P.wall_sd[w][sp] = std::sqrt(sys.wall_t[a][side] / sys.masses[sp]);
                }
            }
        P.gravity[a] = core::gravity_kick(sys.dt, sys.gravity[a]);
    }
    // Stochastic boundaries: Threefry keyed by the system seed (three.rs:232-265); every walled
    // axis and the bath book their exchanges in the ledger.
    P.seed_lo = (uint32_t)sys.seed;
    P.seed_hi = (uint32_t)(sys.seed >> 32);
    P.dt = sys.dt;
    for (uint32_t sp = 0; sp < 4; sp++) P.mass[sp] = sp < sys.nspecies ? sys.masses[sp] : 1.0;
    if (sys.has_bath) {
        P.bath = 1u;
        P.bath_species = sys.bath_species;
        P.bath_p = sys.bath_rate * sys.dt;
        for (uint32_t sp = 0; sp < sys.nspecies; sp++) P.bath_sd[sp] = std::sqrt(sys.bath_t / sys.masses[sp]);
    }
    P.ledger = P.bath || (P.periodic != (1u << D) - 1u) ? 1u : 0u;
    P.mbits = 32 - __builtin_clz(std::max(ncmax - 1, 1u));
    RH_REQUIRE((D == 3 && P.mbits <= 10) || (D == 2 && P.mbits <= 15), "too many cells for Morton keys");
    key_bits_ = D * P.mbits;
    ncell_keys_ = uint64_t(1) << key_bits_;
    for (uint32_t s = 0; s < 4; s++) P.kick_c[s] = s < sys.nspecies ? (float)core::kick_coeff(sys.dt, sys.masses[s], (sys.pinned >> s) & 1u) : 0.0f;
    double hs = 0.5 * skin * (1.0 - 1e-3);
    P.half_skin2 = (float)(hs * hs);
    for (uint32_t a = 0; a < 4; a++)
        for (uint32_t bb = 0; bb < 4; bb++) {
            int k = a * 4 + bb;
            if (a >= sys.nspecies || bb >= sys.nspecies) continue;
            const rh_pair& p = sys.pairs[k];
            RH_REQUIRE(p.rc < 7.9, "cutoff must be below 8 sigma");
            // Pair::gpu_constants (pair.rs:53-56): f64 constants rounded once to f32.
            P.pt.rc2[k] = (float)(p.rc * p.rc);
            P.pt.eps4[k] = (float)(4.0 * p.eps);
            P.pt.sigma2[k] = (float)(p.sigma * p.sigma);
            P.pt.shift[k] = (float)core::pair_shift(p);
            P.pt.six_eps4[k] = 6.0f * P.pt.eps4[k];
            double r = (p.rc + skin) * (1.0 + 1e-4);
            P.pt.list_r2[k] = (float)(r * r);
            P.pt.r2_fast[k] = fast_division_floor(P.pt.rc2[k], P.pt.sigma2[k], P.pt.six_eps4[k]);
            P.pt.eps4_64[k] = 4.0 * p.eps;
            P.pt.sigma2_64[k] = p.sigma * p.sigma;
            P.pt.rc2_64[k] = p.rc * p.rc;
            P.pt.shift_64[k] = core::pair_shift(p);
        }
    for (uint32_t s = 0; s < 4; s++) P.pt.kick_c[s] = P.kick_c[s];
    P.face16 = face_margin16(P.pt, P.half_skin2);
    P.cap = mode.kind == 1 ? mode.cap : 0;
    RH_REQUIRE(P.cap < 65536, "list capacity must be below 65536");
    for (int a = 0; a < D; a++) {
        P.cell_w64[a] = b[a] / (double)P.nc[a];
        P.cell_w[a] = (float)P.cell_w64[a];
    }
    P.block = block_;
    plan_rows();
    if constexpr (D == 2) {
        // md_core::System's boundaries (system.rs:26-61, 204-353)
        const bool has_bonds = !sys.bonds.empty() || sys.nbond_types > 0;
        P.feat = boundary_features2(sys) ? 1u : 0u;
        for (int a = 0; a < 2; a++)
            if (sys.compartments[a]) {
                P.comp_shift[a] = sys.compartments[a];
                P.comp_scale[a] = std::ldexp(1.0, -(int)sys.compartments[a]);
            }
        P.comp_filter = (sys.compartments[0] || sys.compartments[1]) && !sys.diathermal ? 1u : 0u;
        if (sys.has_bath) {
            P.bath_axis = sys.bath_axis;
            P.bath_drift = core::to_step_displacement(sys.dt, sys.bath_drift);
            P.bath_lo = (double)sys.bath_region[0];
            P.bath_hi = (double)sys.bath_region[1];
        }
        if (sys.has_membrane) {
            P.membrane = 1u;
            P.mem_axis = sys.membrane_axis;
            P.mem_blocks = sys.membrane_blocks;
            P.pore_shift = sys.pore_shift;
            P.mem_pos = (double)sys.membrane_position;
            P.pore_mask = sys.pore_shift ? (int64_t(1) << sys.pore_shift) - 1 : 0;
            P.pore_width = sys.pore_width;
        }
        if (sys.has_piston) {
            P.piston = 1u;
            P.piston_free = sys.piston_free;
            P.piston_mass = sys.piston_mass;
            P.piston_load = sys.piston_force;
            P.piston_lo = core::to_fixed(sys.piston_min);
            P.piston_hi = core::to_fixed(sys.piston_max);
        }
        P.bonds = has_bonds ? 1u : 0u;
        for (uint32_t t = 0; t < 4; t++) {
            if (sys.bond_fene[t]) P.bond_fene |= 1u << t;
            P.bond_k[t] = (float)sys.bond_k[t];
            P.bond_r0[t] = (float)sys.bond_r0[t];
        }
    }
    // Listed pairs stay within (list radius + skin) of each other while the list is valid; below
    // 8 sigma the reference's range test cannot fail and is skipped.
    P.range_check = ((rcm + skin) * (1.0 + 1e-4) + skin >= 7.99) ? 1u : 0u;
}

template <int D>
void Engine<D>::upload_tables() {
    RH_HIP(hipMemcpyAsync(d_pt_, &P_.pt, sizeof(PairTable), hipMemcpyHostToDevice, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
    B_.pt = d_pt_;
}

template <int D>
hipEvent_t Engine<D>::timer_event() {
    if (ev_used_ == ev_pool_.size()) {
        hipEvent_t e;
        RH_HIP(hipEventCreate(&e));
        ev_pool_.push_back(e);
    }
    return ev_pool_[ev_used_++];
}

template <int D>
void Engine<D>::record_timer(bool rebuild, hipEvent_t start) {
    hipEvent_t end = timer_event();
    RH_HIP(hipEventRecord(end, stream_));
    timed_.push_back({start, end, rebuild});
}

template <int D>
void Engine<D>::collect_timers() {
    if (timed_.empty()) return;
    RH_HIP(hipStreamSynchronize(stream_));
    for (auto& t : timed_) {
        float ms = 0.f;
        RH_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        if (t.rebuild) {
            rebuild_ms_ += ms;
            rebuild_launches_++;
        } else {
            force_ms_ += ms;
            force_launches_++;
        }
    }
    timed_.clear();
    ev_used_ = 0;
}

// The row build's geometry (3D Verlet mode): rows are the (y, z) cells, binned along x into 8,
// 4, 2 or 1 fine cells per cell, whichever first keeps the extended bins within 2^25.  It needs
// three cells on every axis (distinct rows, a chord shorter than the box), and x below 512
// sigma (x is held absolute in f32 there, within 3e-5 sigma).  A periodic x axis adds ghost
// bins covering the list radius plus one bin at both ends of every row.
template <int D>
void Engine<D>::plan_rows() {
    Params<D>& P = P_;
    P.rows = 0u;
    P.rows_nxf = 0u;
    row_keys_ = 0;
    row_bins_ = 0;
    if constexpr (D == 3) {
        if (mode_.kind != 1) return;
        for (int a = 0; a < 3; a++)
            if (P.nc[a] < 3u) return;
        const double lx = core::to_f64(sys_.box_len[0]);
        if (lx > 512.0) return;
        float lr2max = 0.f;
        for (int k = 0; k < 16; k++) lr2max = std::max(lr2max, P.pt.list_r2[k]);
        const double reach = std::sqrt((double)lr2max) + 1e-3;
        const bool wrap_x = (P.periodic & 1u) != 0u;
        for (uint32_t sub = 8; sub >= 1 && !P.rows; sub /= 2) {
            const uint64_t nxf = (uint64_t)P.nc[0] * sub;
            const double wx = lx / (double)nxf;
            const uint64_t g = wrap_x ? (uint64_t)std::ceil(reach / wx) + 1u : 0u;
            const uint64_t rows = (uint64_t)P.nc[1] * P.nc[2];
            const uint64_t keys = rows * (nxf + 2 * g);
            // a chord never holds a slot and its ghost, and no slot has two ghosts
            if (keys > (uint64_t(1) << 25) || nxf <= 2 * g || !(lx > 2.0 * reach + 2.0 * wx)) continue;
            P.rows = 1u;
            P.rows_nxf = (uint32_t)nxf;
            P.rows_g = (uint32_t)g;
            P.rows_ne = (uint32_t)(nxf + 2 * g);
            row_keys_ = keys;
            row_bins_ = rows * nxf;
        }
        if (!P.rows) return;
        P.rows_xscale = (double)P.rows_nxf / (double)sys_.box_len[0];
        P.rows_inv_wx = (float)((double)P.rows_nxf / lx);
        P.rows_reach = (float)reach;
    }
}

template <int D>
void Engine<D>::alloc_rows() {
    if (P_.rows == 0u || B_.items != nullptr) return;
    // every slot has one copy and at most one ghost; padded, since the build reads past a range
    const size_t items = 2 * n_ + 32u;
    B_.items = (float4*)dalloc(items * sizeof(float4));
    RH_HIP(hipMemset(B_.items, 0, items * sizeof(float4)));
    B_.row_cnt = (uint32_t*)dalloc(row_bins_ * sizeof(uint32_t));
    B_.row_start = (uint32_t*)dalloc((row_keys_ + 1u) * sizeof(uint32_t));
    B_.row_rank = (uint32_t*)dalloc(n_ * sizeof(uint32_t));
    row_scan_temp_ = dalloc(scan_temp_bytes(row_keys_ + 1u));
    key_scan_temp_ = dalloc(scan_temp_bytes(ncell_keys_));
}

template <int D>
void Engine<D>::alloc_list(uint32_t cap) {
    auto release = [&](void* p) {
        if (!p) return;
        auto it = std::find(allocs_.begin(), allocs_.end(), p);
        if (it != allocs_.end()) allocs_.erase(it);
        RH_HIP(hipFree(p));
    };
    if (B_.nbr) alloc_bytes_ -= list_bytes_;
    release(B_.nbr);
    cap = (cap + 3u) & ~3u;  // whole four-entry chunks (the half list stores uint4 chunks)
    P_.cap = cap;
    // The full lists take cap entries per particle; the half list's chunk stream takes cap / 4
    // chunks per particle slot, for every block size up to 4096.
    const size_t slots = (n_ + 4095u) & ~size_t(4095);
    list_bytes_ = (size_t)cap * slots * sizeof(uint32_t);
    B_.nbr = (uint32_t*)dalloc(list_bytes_);
}

template <int D>
Engine<D>::Engine(const SystemDesc& sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, const rh_mode& mode)
    : sys_(sys), mode_(mode), n_(n) {
    RH_REQUIRE(n >= 1, "an engine needs at least one particle");
    RH_REQUIRE(n < (uint64_t(1) << 31), "too many particles for one engine");
    RH_REQUIRE(mode.kind <= 1, "unknown force mode");
    if (mode.kind == 1) RH_REQUIRE(mode.skin >= 0.0 && mode.cap >= 1, "Verlet mode needs skin >= 0 and cap >= 1");
    for (uint64_t i = 0; i < n; i++) RH_REQUIRE(species[i] < sys.nspecies, "species out of range");
    std::memset(&B_, 0, sizeof B_);
    // Large systems default to large blocks (fewer pairs cross blocks); smaller ones keep enough
    // blocks to fill the chip (measured best: 512 below 2^18 particles, 1024 to 2^20).
    block_ = n >= (uint64_t(1) << 20) ? 2048u : (n >= (uint64_t(1) << 18) ? 1024u : 512u);
    build_params(sys, mode);
    RH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (int p = 0; p < 2; p++) {
        for (int a = 0; a < D; a++) {
            B_.pos[p][a] = (double*)dalloc(n * sizeof(double));
            B_.vel[p][a] = (int32_t*)dalloc(n * sizeof(int32_t));
        }
        B_.species[p] = (uint32_t*)dalloc(n * sizeof(uint32_t));
        B_.aos[p] = (double*)dalloc(n * (D == 3 ? 4 : 2) * sizeof(double));
        B_.ids[p] = (uint32_t*)dalloc(n * sizeof(uint32_t));
    }
    for (int a = 0; a < D; a++) {
        B_.ref[a] = (double*)dalloc(n * sizeof(double));
        B_.kicks[a] = (int32_t*)dalloc(n * sizeof(int32_t));
    }
    B_.nbr_count = (uint32_t*)dalloc(n * sizeof(uint32_t));
    {
        // two counts (in-block and cross-block chunks) per 64-particle wave, for every block size
        // up to 4096; waves past the last particle stay empty
        const size_t waves = ((n + 4095u) & ~size_t(4095)) / 64u;
        B_.wave_chunks = (uint32_t*)dalloc(2 * waves * sizeof(uint32_t));
        RH_HIP(hipMemset(B_.wave_chunks, 0, 2 * waves * sizeof(uint32_t)));
    }
    B_.rel = (float4*)dalloc((n + 32u) * sizeof(float4));  // padded: the list build reads up to 31 past a cell
    alloc_rows();
    B_.cell_start = (uint32_t*)dalloc(ncell_keys_ * sizeof(uint32_t));
    B_.cell_end = (uint32_t*)dalloc(ncell_keys_ * sizeof(uint32_t));
    B_.pe = (double*)dalloc(n * sizeof(double));
    B_.ke = (double*)dalloc(n * sizeof(double));
    B_.ctl = (uint32_t*)dalloc(CTL_WORDS * sizeof(uint32_t));
    B_.ledger = (unsigned long long*)dalloc(LEDGER_WORDS * sizeof(unsigned long long));
    RH_HIP(hipMemset(B_.ledger, 0, LEDGER_WORDS * sizeof(unsigned long long)));
    if constexpr (D == 2) {
        if (P_.bonds) {
            for (uint64_t b = 0; b < sys.bonds.size() / 3; b++)
                RH_REQUIRE(sys.bonds[3 * b] < n && sys.bonds[3 * b + 1] < n && sys.bonds[3 * b] != sys.bonds[3 * b + 1], "bond ids must be distinct particles");
            B_.slot_of = (uint32_t*)dalloc(n * sizeof(uint32_t));
            B_.partners = (const uint32_t*)dalloc(n * 8 * sizeof(uint32_t));
            write_partners();
        }
        if (P_.piston) {
            B_.piston = (PistonDev*)dalloc(sizeof(PistonDev));
            PistonDev p{};
            p.w = core::to_fixed(sys.piston_start);
            p.v = sys.piston_free ? 0 : core::to_step_displacement(sys.dt, sys.piston_speed);
            RH_HIP(hipMemcpy(B_.piston, &p, sizeof p, hipMemcpyHostToDevice));
        }
        if (P_.feat) alloc_boundary_buffers();
    }
    reduce_temp_ = dalloc(reduce_temp_bytes(n));
    d_scalar_ = (double*)dalloc(64);
    if (mode.kind == 1) alloc_list(mode.cap);
    sort_.reserve(n);
    sort_bytes_ = sort_.bytes();
    constexpr int NSLOTS = 8;
    RH_HIP(hipHostMalloc((void**)&h_ctl_, NSLOTS * CTL_WORDS * sizeof(uint32_t), hipHostMallocDefault));
    ev_.resize(NSLOTS);
    for (auto& e : ev_) RH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    d_pt_ = (PairTable*)dalloc(sizeof(PairTable));
    upload_tables();
    since_sort_ = sort_interval_;
    upload(pos, vel, species);
}

template <int D>
Engine<D>::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) (void)hipFree(p);
    for (auto& e : ev_) (void)hipEventDestroy(e);
    for (auto& e : ev_pool_) (void)hipEventDestroy(e);
    if (h_ctl_) (void)hipHostFree(h_ctl_);
    sort_.release();
    if (stream_) (void)hipStreamDestroy(stream_);
}

template <int D>
void Engine<D>::write_ctl(uint32_t word, uint32_t value) {
    k_set_word<<<1, 1, 0, stream_>>>(B_.ctl, word, value);
    RH_HIP(hipGetLastError());
}

template <int D>
void Engine<D>::read_ctl(uint32_t* out) {
    RH_HIP(hipMemcpyAsync(h_ctl_, B_.ctl, CTL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
    std::memcpy(out, h_ctl_, CTL_WORDS * sizeof(uint32_t));
}

template <int D>
void Engine<D>::upload(const int64_t* pos, const int32_t* vel, const uint32_t* species) {
    RH_HIP(hipStreamSynchronize(stream_));
    for (uint64_t i = 0; i < n_; i++) {
        RH_REQUIRE(species[i] < P_.nspecies, "species out of range");
        for (int a = 0; a < D; a++)
            RH_REQUIRE(pos[D * i + a] >= 0 && pos[D * i + a] < P_.box[a], "position outside the box");
    }
    std::vector<double> px(n_);
    std::vector<int32_t> vx(n_);
    std::vector<uint32_t> ids(n_);
    for (int a = 0; a < D; a++) {
        for (uint64_t i = 0; i < n_; i++) px[i] = (double)pos[D * i + a];  // exact: |x| < 2^52
        RH_HIP(hipMemcpy(B_.pos[0][a], px.data(), n_ * sizeof(double), hipMemcpyHostToDevice));
        for (uint64_t i = 0; i < n_; i++) vx[i] = vel[D * i + a];
        RH_HIP(hipMemcpy(B_.vel[0][a], vx.data(), n_ * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    RH_HIP(hipMemcpy(B_.species[0], species, n_ * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (uint64_t i = 0; i < n_; i++) ids[i] = (uint32_t)i;
    RH_HIP(hipMemcpy(B_.ids[0], ids.data(), n_ * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (B_.slot_of) RH_HIP(hipMemcpy(B_.slot_of, ids.data(), n_ * sizeof(uint32_t), hipMemcpyHostToDevice));
    std::vector<uint32_t> ctl(CTL_WORDS, 0u);
    ctl[CTL_STALE] = 1u;
    RH_HIP(hipMemcpy(B_.ctl, ctl.data(), CTL_WORDS * sizeof(uint32_t), hipMemcpyHostToDevice));
    steps_ = 0;
    since_sort_ = sort_interval_;
}

template <int D>
void Engine<D>::enqueue_cells_at_current(const uint32_t* gate) {
    uint32_t nb = blocks_for(n_, TPB);
    k_keys<D><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.keys[0], sort_.vals[0], gate);
    RH_HIP(hipGetLastError());
    int r = sort_.sort(n_, key_bits_, gate, stream_);
    fill_u32(B_.cell_start, 0u, ncell_keys_, gate, stream_);
    fill_u32(B_.cell_end, 0u, ncell_keys_, gate, stream_);
    k_ranges<<<nb, TPB, 0, stream_>>>(sort_.keys[r], (uint32_t)n_, B_.cell_start, B_.cell_end, gate);
    RH_HIP(hipGetLastError());
}

template <int D>
void Engine<D>::enqueue_force_cells(const uint32_t* gate) {
    int r = RadixSort::result_index(key_bits_);
    uint32_t nb = blocks_for(n_, TPB);
    if (P_.nspecies > 1) k_force_cells<D, true><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.vals[r], gate);
    else k_force_cells<D, false><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.vals[r], gate);
    RH_HIP(hipGetLastError());
}

// The gated rebuild chain: keys, sort, reorder, cell ranges of the new order, list build.
template <int D>
void Engine<D>::enqueue_rebuild(bool gated) {
    uint32_t* gate = B_.ctl + CTL_DO_REBUILD;
    if (gated) k_rebuild_decide<<<1, 1, 0, stream_>>>(B_.ctl);
    else k_rebuild_force<<<1, 1, 0, stream_>>>(B_.ctl);
    RH_HIP(hipGetLastError());
    hipEvent_t t0 = nullptr;
    if (timing_) {
        t0 = timer_event();
        RH_HIP(hipEventRecord(t0, stream_));
    }
    uint32_t nb = blocks_for(n_, TPB);
    const bool rows = mode_.kind == 1 && row_path();
    if (rows) {
        // counting sort of the cell keys (the cell-start array holds the counts, then the first
        // slot of each key), and the reorder
        fill_u32(B_.cell_start, 0u, ncell_keys_, gate, stream_);
        k_keys_count<D><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.keys[0], sort_.vals[0], B_.cell_start, gate);
        RH_HIP(hipGetLastError());
        exclusive_scan_u32(B_.cell_start, B_.cell_start, ncell_keys_, key_scan_temp_, gate, stream_);
        k_keys_place<<<nb, TPB, 0, stream_>>>((uint32_t)n_, sort_.keys[0], sort_.vals[0], B_.cell_start, sort_.vals[1], gate);
        RH_HIP(hipGetLastError());
        k_gather<D><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.vals[1], 1, gate);
        RH_HIP(hipGetLastError());
    } else {
        k_keys<D><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.keys[0], sort_.vals[0], gate);
        RH_HIP(hipGetLastError());
        int r = sort_.sort(n_, key_bits_, gate, stream_);
        k_gather<D><<<nb, TPB, 0, stream_>>>(P_, B_, sort_.vals[r], 1, gate);
        RH_HIP(hipGetLastError());
        // the cell ranges of the new order, for the stencil and full-list builds (readouts that
        // need the ranges recompute them)
        fill_u32(B_.cell_start, 0u, ncell_keys_, gate, stream_);
        fill_u32(B_.cell_end, 0u, ncell_keys_, gate, stream_);
        k_ranges<<<nb, TPB, 0, stream_>>>(sort_.keys[r], (uint32_t)n_, B_.cell_start, B_.cell_end, gate);
        RH_HIP(hipGetLastError());
    }
    if (mode_.kind == 1) {
        bool multi = P_.nspecies > 1;
        if constexpr (D == 3) {
            if (rows) {
                fill_u32(B_.row_cnt, 0u, row_bins_, gate, stream_);
                k_row_count<<<nb, TPB, 0, stream_>>>(P_, B_, gate);
                RH_HIP(hipGetLastError());
                k_row_ext<<<blocks_for(row_keys_ + 1u, TPB), TPB, 0, stream_>>>(P_, B_, (uint32_t)row_keys_, gate);
                RH_HIP(hipGetLastError());
                exclusive_scan_u32(B_.row_start, B_.row_start, row_keys_ + 1u, row_scan_temp_, gate, stream_);
                k_row_scatter<<<nb, TPB, 0, stream_>>>(P_, B_, gate);
                RH_HIP(hipGetLastError());
                if (multi) k_build_rows<true><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
                else k_build_rows<false><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
            }
        }
        if (row_path()) {
            // built above
        } else if (half_path()) {
            bool wrap = false;
            for (int a = 0; a < D; a++) wrap |= P_.nc[a] == 1u && ((P_.periodic >> a) & 1u);
            if (wrap) {
                if (multi) k_build_half<D, true, true><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
                else k_build_half<D, false, true><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
            } else {
                if (multi) k_build_half<D, true, false><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
                else k_build_half<D, false, false><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
            }
        } else {
            if (multi) k_build_list<D, true><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
            else k_build_list<D, false><<<nb, TPB, 0, stream_>>>(P_, B_, gate);
        }
        RH_HIP(hipGetLastError());
    }
    if constexpr (D == 2) {
        if (B_.slot_of) {
            k_slot_of<<<nb, TPB, 0, stream_>>>(P_, B_, gate);
            RH_HIP(hipGetLastError());
        }
    }
    if (timing_) record_timer(true, t0);
    k_rebuild_commit<<<1, 1, 0, stream_>>>(B_.ctl);
    RH_HIP(hipGetLastError());
}

template <int D>
void Engine<D>::handle_overflow() {
    uint32_t ctl[CTL_WORDS];
    read_ctl(ctl);
    uint32_t want = std::max<uint32_t>(ctl[CTL_MAXCOUNT] + ctl[CTL_MAXCOUNT] / 4 + 8, P_.cap + P_.cap / 2);
    alloc_list(want);
    write_ctl(CTL_OVERFLOW, 0u);
    write_ctl(CTL_STALE, 1u);
}

template <int D>
void Engine<D>::step_list(uint64_t steps) {
    constexpr int NSLOTS = 8, DEPTH = 3;
    uint64_t left = steps;
    while (left > 0) {
        uint32_t add = (uint32_t)std::min<uint64_t>(left, 1u << 30);
        left -= add;
        k_add_remaining<<<1, 1, 0, stream_>>>(B_.ctl, add);
        RH_HIP(hipGetLastError());
        struct Flight {
            int slot;
            uint32_t k;
        };
        std::deque<Flight> flight;
        uint64_t remaining = add;
        uint32_t last_done = 0, last_rebuilds = 0;
        bool have_last = false;
        int next_slot = 0;
        bool done = false;
        while (!done) {
            // Chunk length: a little over the expected list lifetime, no more than is left.
            double expect_inflight = 0.0;
            for (auto& f : flight) expect_inflight += std::min<double>(f.k, lifetime_);
            double want = std::ceil(lifetime_ * 1.25) + 1.0;
            double room = (double)remaining - expect_inflight;
            if (room >= 1.0) {
                uint32_t k = (uint32_t)std::max(1.0, std::min({want, room + 1.0, 512.0}));
                enqueue_rebuild(true);
                for (uint32_t s = 0; s < k; s++) {
                    hipEvent_t t0 = nullptr;
                    if (timing_) {
                        t0 = timer_event();
                        RH_HIP(hipEventRecord(t0, stream_));
                    }
                    launch_step();
                    if (timing_) record_timer(false, t0);
                    k_step_commit<<<1, 1, 0, stream_>>>(B_.ctl);
                    RH_HIP(hipGetLastError());
                }
                int slot = next_slot++ % NSLOTS;
                RH_HIP(hipMemcpyAsync(h_ctl_ + slot * CTL_WORDS, B_.ctl, CTL_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
                RH_HIP(hipEventRecord(ev_[slot], stream_));
                flight.push_back({slot, k});
            }
            bool must_wait = flight.size() >= (size_t)DEPTH || room < 1.0;
            while (!flight.empty() && (must_wait || hipEventQuery(ev_[flight.front().slot]) == hipSuccess)) {
                must_wait = false;
                Flight f = flight.front();
                flight.pop_front();
                RH_HIP(hipEventSynchronize(ev_[f.slot]));
                const uint32_t* c = h_ctl_ + f.slot * CTL_WORDS;
                remaining = c[CTL_REMAINING];
                uint32_t dn = c[CTL_DONE_LO], rb = c[CTL_REBUILDS];
                if (have_last && rb > last_rebuilds) {
                    double life = (double)(dn - last_done) / (double)(rb - last_rebuilds);
                    lifetime_ = 0.5 * lifetime_ + 0.5 * std::max(1.0, life);
                } else if (have_last && rb == last_rebuilds && dn > last_done) {
                    lifetime_ = std::min(512.0, std::max(lifetime_, (double)(dn - last_done) * 1.5));
                }
                last_done = dn;
                last_rebuilds = rb;
                have_last = true;
                if (c[CTL_OVERFLOW]) {
                    RH_HIP(hipStreamSynchronize(stream_));
                    flight.clear();
                    handle_overflow();
                    uint32_t now[CTL_WORDS];
                    read_ctl(now);
                    remaining = now[CTL_REMAINING];
                    have_last = false;
                    break;
                }
                if (remaining == 0) {
                    done = true;
                    break;
                }
            }
        }
    }
}

template <int D>
void Engine<D>::launch_step() {
    if constexpr (D == 2) {
        if (split_path()) {
            launch_boundary_step(0u, true);
            return;
        }
    }
    const bool multi = P_.nspecies > 1;
    if (!half_path()) {
        uint32_t nb = blocks_for(n_, TPB);
        if (P_.ledger) {
            if (multi) k_step_list<D, true, true><<<nb, TPB, 0, stream_>>>(P_, B_);
            else k_step_list<D, false, true><<<nb, TPB, 0, stream_>>>(P_, B_);
        } else {
            if (multi) k_step_list<D, true, false><<<nb, TPB, 0, stream_>>>(P_, B_);
            else k_step_list<D, false, false><<<nb, TPB, 0, stream_>>>(P_, B_);
        }
    } else {
        uint32_t bs = P_.block;
        uint32_t nb = blocks_for(n_, bs);
#define RH_STEP_HALF(BS)                                                                   \
    if (P_.ledger) {                                                                        \
        if (multi) k_step_half<D, true, BS, true><<<nb, nt, 0, stream_>>>(P_, B_);          \
        else k_step_half<D, false, BS, true><<<nb, nt, 0, stream_>>>(P_, B_);               \
    } else {                                                                                \
        if (multi) k_step_half<D, true, BS, false><<<nb, nt, 0, stream_>>>(P_, B_);         \
        else k_step_half<D, false, BS, false><<<nb, nt, 0, stream_>>>(P_, B_);              \
    }
        const uint32_t nt = bs > 1024u ? 1024u : bs;
        if (bs == 256) {
            RH_STEP_HALF(256)
        } else if (bs == 512) {
            RH_STEP_HALF(512)
        } else if (bs == 1024) {
            RH_STEP_HALF(1024)
        } else if (bs == 2048) {
            RH_STEP_HALF(2048)
        } else {
            RH_STEP_HALF(4096)
        }
#undef RH_STEP_HALF
    }
    RH_HIP(hipGetLastError());
}

template <int D>
void Engine<D>::tune(int variant, uint32_t block) {
    RH_REQUIRE(variant >= 0 && variant <= 2, "kernel variant must be 0 (full list), 1 (half list) or 2 (half list, cell-stencil build)");
    RH_REQUIRE(block == 256 || block == 512 || block == 1024 || block == 2048 || block == 4096, "block must be 256, 512, 1024, 2048 or 4096");
    RH_HIP(hipStreamSynchronize(stream_));
    variant_ = variant;
    block_ = block;
    P_.block = block;
    write_ctl(CTL_STALE, 1u);
}

template <int D>
void Engine<D>::step_cells(uint64_t steps) {
    uint32_t nb = blocks_for(n_, TPB);
    uint32_t& since = since_sort_;
    for (uint64_t s = 0; s < steps; s++) {
        if (sort_interval_ > 0 && since >= sort_interval_) {
            // Morton reorder of storage (engine3.rs:502-507), then forces on the sorted slots.
            enqueue_rebuild(false);
            since = 0;
        }
        enqueue_cells_at_current(nullptr);
        enqueue_force_cells(nullptr);
        if (P_.feat) {
            launch_boundary_step((uint32_t)s, false);
        } else {
            k_integrate<D><<<nb, TPB, 0, stream_>>>(P_, B_, (uint32_t)s);
            RH_HIP(hipGetLastError());
        }
        since++;
    }
    k_cells_commit<<<1, 1, 0, stream_>>>(B_.ctl, (uint32_t)steps);
    RH_HIP(hipGetLastError());
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
void Engine<D>::step(uint64_t steps) {
    if (steps == 0) return;
    if (mode_.kind == 1) step_list(steps);
    else step_cells(steps);
    collect_timers();
    steps_ += (int64_t)steps;
}

template <int D>
void Engine<D>::kicks(int32_t* kicks, double* potential) {
    enqueue_cells_at_current(nullptr);
    enqueue_force_cells(nullptr);
    if (P_.feat) enqueue_extras(EX_PE);
    uint32_t ctl[CTL_WORDS];
    read_ctl(ctl);
    uint32_t pm = ctl[CTL_PMETA];
    std::vector<int32_t> k(n_);
    std::vector<uint32_t> ids(n_);
    std::vector<double> pe(n_);
    RH_HIP(hipMemcpy(ids.data(), B_.ids[pm], n_ * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int a = 0; a < D; a++) {
        RH_HIP(hipMemcpy(k.data(), B_.kicks[a], n_ * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (kicks)
            for (uint64_t s = 0; s < n_; s++) kicks[D * (uint64_t)ids[s] + a] = k[s];
    }
    RH_HIP(hipMemcpy(pe.data(), B_.pe, n_ * sizeof(double), hipMemcpyDeviceToHost));
    double total = 0.0;
    std::vector<double> by_id(n_);
    for (uint64_t s = 0; s < n_; s++) by_id[ids[s]] = pe[s];
    for (uint64_t i = 0; i < n_; i++) total += by_id[i];
    if (potential) *potential = total;
}

template <int D>
double Engine<D>::potential() {
    double pe = 0.0;
    kicks(nullptr, &pe);
    return pe;
}

template <int D>
void Engine<D>::reverse() {
    enqueue_cells_at_current(nullptr);
    enqueue_force_cells(nullptr);
    if (P_.feat) enqueue_extras(EX_PISTON);
    k_reverse<D><<<blocks_for(n_, TPB), TPB, 0, stream_>>>(P_, B_);
    RH_HIP(hipGetLastError());
    if constexpr (D == 2) {
        if (P_.piston) {
            k_piston_reverse2<<<1, 1, 0, stream_>>>(P_, B_);
            RH_HIP(hipGetLastError());
        }
    }
    RH_HIP(hipStreamSynchronize(stream_));
    steps_ = -steps_;
}

template <int D>
void Engine<D>::read_state(int64_t* pos, int32_t* vel, uint32_t* species) {
    uint32_t ctl[CTL_WORDS];
    read_ctl(ctl);
    uint32_t pp = ctl[CTL_PPOS], pm = ctl[CTL_PMETA];
    std::vector<uint32_t> ids(n_), sp(n_);
    std::vector<double> px(n_);
    std::vector<int32_t> vx(n_);
    RH_HIP(hipMemcpy(ids.data(), B_.ids[pm], n_ * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int a = 0; a < D; a++) {
        RH_HIP(hipMemcpy(px.data(), B_.pos[pp][a], n_ * sizeof(double), hipMemcpyDeviceToHost));
        RH_HIP(hipMemcpy(vx.data(), B_.vel[pm][a], n_ * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (uint64_t s = 0; s < n_; s++) {
            if (pos) pos[D * (uint64_t)ids[s] + a] = (int64_t)px[s];
            if (vel) vel[D * (uint64_t)ids[s] + a] = vx[s];
        }
    }
    if (species) {
        RH_HIP(hipMemcpy(sp.data(), B_.species[pm], n_ * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint64_t s = 0; s < n_; s++) species[ids[s]] = sp[s];
    }
}

// Every slot's kinetic energy from its half-step velocity, and their sum (GpuEngine3::
// measure_kinetic, engine3.rs:618-624), enqueued on the engine's stream.
template <int D>
void Engine<D>::measure_kinetic() {
    k_kinetic<D><<<blocks_for(n_, TPB), TPB, 0, stream_>>>(P_, B_, sys_.masses[0], sys_.masses[1], sys_.masses[2], sys_.masses[3], sys_.dt);
    RH_HIP(hipGetLastError());
    reduce_sum_f64(B_.ke, d_scalar_, n_, reduce_temp_, stream_);
}

// Kinetic energies (f32, as the reference returns them; computed in f64 here) and species in
// particle-id order (engine3.rs:626-631).
template <int D>
void Engine<D>::read_kinetic(float* ke, uint32_t* species) {
    measure_kinetic();
    uint32_t ctl[CTL_WORDS];
    read_ctl(ctl);
    const uint32_t pm = ctl[CTL_PMETA];
    std::vector<uint32_t> ids(n_), sp(n_);
    std::vector<double> e(n_);
    RH_HIP(hipMemcpy(ids.data(), B_.ids[pm], n_ * sizeof(uint32_t), hipMemcpyDeviceToHost));
    RH_HIP(hipMemcpy(e.data(), B_.ke, n_ * sizeof(double), hipMemcpyDeviceToHost));
    RH_HIP(hipMemcpy(sp.data(), B_.species[pm], n_ * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint64_t s = 0; s < n_; s++) {
        if (ke) ke[ids[s]] = (float)e[s];
        if (species) species[ids[s]] = sp[s];
    }
}

// A force evaluation at the current positions without moving (GpuEngine::compute_forces,
// engine.rs:875-878); kicks() reads its result.
template <int D>
void Engine<D>::compute_forces() {
    enqueue_cells_at_current(nullptr);
    enqueue_force_cells(nullptr);
    if (P_.feat) enqueue_extras(EX_PE);
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
double Engine<D>::kinetic_total() {
    measure_kinetic();
    double out = 0.0;
    RH_HIP(hipMemcpyAsync(&out, d_scalar_, sizeof(double), hipMemcpyDeviceToHost, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
    return out;
}

template <int D>
void Engine<D>::update_system(const SystemDesc& sys) {
    // engine3.rs:536-559: same box, species count and wall layout; cutoffs may not grow.
    for (int a = 0; a < D; a++) {
        RH_REQUIRE(sys.box_len[a] == sys_.box_len[a], "update_system: the box must stay the same");
        RH_REQUIRE((sys.walled[a] != 0) == (sys_.walled[a] != 0), "update_system: walls cannot be added or removed");
    }
    RH_REQUIRE(sys.nspecies == sys_.nspecies, "update_system: the species count must stay the same");
    RH_REQUIRE(core::rc_max(sys.pairs, sys.nspecies) <= core::rc_max(sys_.pairs, sys_.nspecies) + 1e-12, "cutoffs cannot grow at run time");
    RH_REQUIRE(sys.has_bath == sys_.has_bath, "the bath can be retuned, not added or removed");
    if constexpr (D == 2) {
        // engine.rs:895-906
        RH_REQUIRE(sys.compartments[0] == sys_.compartments[0] && sys.compartments[1] == sys_.compartments[1], "update_system: compartments cannot change");
        RH_REQUIRE(sys.bonds == sys_.bonds && (sys.nbond_types > 0) == (sys_.nbond_types > 0), "bonds cannot change at run time");
        RH_REQUIRE(sys.has_piston == sys_.has_piston && (!sys.has_piston || sys.piston_free == sys_.piston_free), "update_system: the piston and its drive cannot change");
        RH_REQUIRE(P_.arith == 0u || std::memcmp(sys.pairs, sys_.pairs, sizeof sys.pairs) == 0, "integer tables are fixed at creation");
        RH_REQUIRE(P_.arith == 0u || (!boundary_features2(sys) && !has_soft_walls(sys)), "integer forces cover pair forces only");
    }
    Params<D> old = P_;
    uint32_t old_bits = key_bits_;
    uint64_t old_keys = ncell_keys_, old_row_keys = row_keys_, old_row_bins = row_bins_;
    build_params(sys, mode_);
    key_bits_ = old_bits;
    ncell_keys_ = old_keys;
    row_keys_ = old_row_keys;
    row_bins_ = old_row_bins;
    // Keep the cell geometry the buffers were sized for; a pair whose list radius grew is listed
    // further out from the next rebuild on.
    for (int a = 0; a < 3; a++) {
        P_.nc[a] = old.nc[a];
        P_.cell_scale[a] = old.cell_scale[a];
        P_.cell_w[a] = old.cell_w[a];
        P_.cell_w64[a] = old.cell_w64[a];
    }
    P_.mbits = old.mbits;
    P_.cap = old.cap;
    // the row build's geometry and reach follow the kept cells and list radii
    P_.rows = old.rows;
    P_.rows_nxf = old.rows_nxf;
    P_.rows_g = old.rows_g;
    P_.rows_ne = old.rows_ne;
    P_.rows_xscale = old.rows_xscale;
    P_.rows_inv_wx = old.rows_inv_wx;
    P_.rows_reach = old.rows_reach;
    P_.half_skin2 = old.half_skin2;
    // Cutoffs only shrink, so a range test that was needed stays safe; keeping it keeps the list
    // in the form it was built in (full lists for the range-tested kernel).
    P_.range_check = old.range_check;
    bool grew = false;
    for (int k = 0; k < 16; k++) {
        grew |= P_.pt.list_r2[k] > old.pt.list_r2[k];
        P_.pt.list_r2[k] = std::max(P_.pt.list_r2[k], old.pt.list_r2[k]);
    }
    P_.face16 = face_margin16(P_.pt, P_.half_skin2);  // the kept list radii and skin
    if (grew) write_ctl(CTL_STALE, 1u);
    P_.arith = old.arith;
    sys_ = sys;
    if (P_.feat && !old.feat) alloc_boundary_buffers();
    // The boundary path reads a full list, the plain path the half list: a change of path needs
    // the list rebuilt in the other form before the next step.
    if (P_.feat != old.feat) write_ctl(CTL_STALE, 1u);
    if (P_.arith) upload_int_tables();  // the tables follow the new pairs, masses and dt
    upload_tables();
}

template <int D>
void Engine<D>::kick_region(const double lo[3], const double hi[3], double scale, const double add[3]) {
    int32_t a[3] = {0, 0, 0};
    for (int k = 0; k < D; k++) a[k] = core::to_step_displacement(sys_.dt, add[k]);
    k_kick_region<D><<<blocks_for(n_, TPB), TPB, 0, stream_>>>(P_, B_, (float)lo[0], (float)lo[1], D == 3 ? (float)lo[2] : 0.f, (float)hi[0], (float)hi[1],
                                                                D == 3 ? (float)hi[2] : 0.f, (float)scale, a[0], a[1], a[2]);
    RH_HIP(hipGetLastError());
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
void Engine<D>::set_species(const uint32_t* ids, uint64_t count, uint32_t species) {
    RH_REQUIRE(species < P_.nspecies, "species out of range");
    std::vector<uint32_t> list(ids, ids + count);
    std::sort(list.begin(), list.end());
    list.erase(std::unique(list.begin(), list.end()), list.end());
    if (list.empty()) return;
    if (list.size() > set_list_cap_) {
        if (d_set_list_) {
            auto it = std::find(allocs_.begin(), allocs_.end(), (void*)d_set_list_);
            if (it != allocs_.end()) allocs_.erase(it);
            RH_HIP(hipFree(d_set_list_));
        }
        d_set_list_ = (uint32_t*)dalloc(list.size() * sizeof(uint32_t));
        set_list_cap_ = list.size();
    }
    RH_HIP(hipMemcpyAsync(d_set_list_, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
    k_set_species<D><<<blocks_for(n_, TPB), TPB, 0, stream_>>>(P_, B_, d_set_list_, (uint32_t)list.size(), species);
    RH_HIP(hipGetLastError());
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
void Engine<D>::set_species_velocity(uint32_t species, const double v[3]) {
    int32_t a[3] = {0, 0, 0};
    for (int k = 0; k < D; k++) a[k] = core::to_step_displacement(sys_.dt, v[k]);
    k_set_species_velocity<D><<<blocks_for(n_, TPB), TPB, 0, stream_>>>(P_, B_, species, a[0], a[1], a[2]);
    RH_HIP(hipGetLastError());
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
void Engine<D>::upload_int_tables() {
    const core::IntTables t = core::int_tables(sys_.pairs, sys_.nspecies, sys_.masses, sys_.pinned, sys_.dt);
    RH_HIP(hipStreamSynchronize(stream_));
    int32_t* tab = (int32_t*)dalloc(t.data.size() * sizeof(int32_t));
    uint32_t* rc2 = (uint32_t*)dalloc(sizeof t.rc2);
    RH_HIP(hipMemcpy(tab, t.data.data(), t.data.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    RH_HIP(hipMemcpy(rc2, t.rc2, sizeof t.rc2, hipMemcpyHostToDevice));
    B_.itab = tab;
    B_.irc2 = rc2;
    P_.itab_len = (uint32_t)t.len;
}

// md::Arith (engine.rs:92-100, with_options :609-625): Int replaces the f32 pair forces by the
// all-integer tables, which give bit-identical trajectories on every device.  Pair forces only.
template <int D>
void Engine<D>::set_arith(int arith) {
    RH_REQUIRE(arith == 0 || arith == 1, "arith must be 0 (f32 pair forces) or 1 (integer tables)");
    if (arith == 1) {
        RH_REQUIRE(D == 2, "integer force tables are a 2D feature");
        RH_REQUIRE(P_.soft_walls == 0 && !P_.feat, "integer forces cover pair forces only");
        upload_int_tables();
    }
    RH_HIP(hipStreamSynchronize(stream_));
    const bool was_split = split_path();
    P_.arith = (uint32_t)arith;
    if (split_path() && !was_split) alloc_boundary_buffers();
    // Verlet mode: integer forces run over the full list of the split path
    if (mode_.kind == 1 && split_path() != was_split) write_ctl(CTL_STALE, 1u);
}

// Ledger3 (engine3.rs:199-205): energy and impulse exchanged with each wall (2 axis + side)
// and the bath (6) since the engine was made, and the hits.
template <int D>
void Engine<D>::read_ledger(rh_ledger3* out) {
    unsigned long long L[LEDGER_WORDS];
    RH_HIP(hipMemcpyAsync(L, B_.ledger, sizeof L, hipMemcpyDeviceToHost, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
    std::memset(out, 0, sizeof *out);
    for (int w = 0; w < 7; w++) {
        out->energy[w] = (double)(long long)L[3 * w] / LEDGER_ONE;
        out->impulse[w] = (double)(long long)L[3 * w + 1] / LEDGER_ONE;
        out->hits[w] = L[3 * w + 2];
    }
}

template <int D>
rh_list_stats Engine<D>::list_stats() {
    uint32_t ctl[CTL_WORDS];
    read_ctl(ctl);
    rh_list_stats s;
    s.rebuilds = ctl[CTL_REBUILDS];
    s.overflowed = ctl[CTL_OVERFLOW];
    s.cap = P_.cap;
    s.max_count = ctl[CTL_MAXCOUNT];
    s.steps_done = ((uint64_t)ctl[CTL_DONE_HI] << 32) | ctl[CTL_DONE_LO];
    s.launches_skipped = ctl[CTL_SKIPPED];
    return s;
}

template <int D>
void Engine<D>::kernel_ms(double* force_ms, uint64_t* force_launches, double* rebuild_ms, uint64_t* rebuilds) const {
    if (force_ms) *force_ms = force_ms_;
    if (force_launches) *force_launches = force_launches_;
    if (rebuild_ms) *rebuild_ms = rebuild_ms_;
    if (rebuilds) *rebuilds = rebuild_launches_;
}

// ---- the 2D boundary path (md::GpuEngine, engine.rs:567-1098) ----

// Entering the boundary path (at creation, or when update_system adds heat-bath walls, a
// membrane ...): the per-compartment ledger starts from the totals.  Until now only the totals
// were booked, and a system off the boundary path has no compartments (they cannot change at run
// time), so every exchange so far belongs to compartment 0 along its wall.
template <int D>
void Engine<D>::alloc_boundary_buffers() {
    const size_t words = 7 * LEDGER2_COMPARTMENTS * 3;
    if (!B_.cledger) B_.cledger = (unsigned long long*)dalloc(words * sizeof(unsigned long long));
    std::vector<unsigned long long> init(words, 0ull);
    RH_HIP(hipStreamSynchronize(stream_));
    unsigned long long tot[LEDGER_WORDS];
    RH_HIP(hipMemcpy(tot, B_.ledger, sizeof tot, hipMemcpyDeviceToHost));
    for (int w = 0; w < 7; w++)
        for (int f = 0; f < 3; f++) init[3 * (w * LEDGER2_COMPARTMENTS) + f] = tot[3 * w + f];
    RH_HIP(hipMemcpy(B_.cledger, init.data(), words * sizeof(unsigned long long), hipMemcpyHostToDevice));
}

// System::partner_table (system.rs:312-324): per particle id, up to 4 (partner id, bond type).
template <int D>
void Engine<D>::write_partners() {
    std::vector<uint32_t> t(n_ * 8, UINT32_MAX);
    std::vector<uint8_t> used(n_, 0);
    const auto& bonds = sys_.bonds;
    for (size_t b = 0; b + 2 < bonds.size(); b += 3) {
        const uint32_t ends[2][2] = {{bonds[b], bonds[b + 1]}, {bonds[b + 1], bonds[b]}};
        for (const auto& e : ends) {
            const uint32_t me = e[0];
            RH_REQUIRE(used[me] < 4, "a particle has more than 4 bonds");
            t[8 * (size_t)me + 2 * used[me]] = e[1];
            t[8 * (size_t)me + 2 * used[me] + 1] = bonds[b + 2];
            used[me]++;
        }
    }
    RH_HIP(hipStreamSynchronize(stream_));
    if (partners_h_.empty()) {
        RH_HIP(hipMemcpy((void*)B_.partners, t.data(), t.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    } else {
        // only the rows that changed
        for (uint64_t id = 0; id < n_; id++)
            if (std::memcmp(&t[8 * id], &partners_h_[8 * id], 8 * sizeof(uint32_t)) != 0)
                RH_HIP(hipMemcpy((void*)(B_.partners + 8 * id), &t[8 * id], 8 * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    partners_h_.swap(t);
}

template <int D>
void Engine<D>::enqueue_extras(uint32_t flags) {
    if constexpr (D == 2) {
        k_extras2<<<blocks_for(n_, TPB), TPB, 0, stream_>>>(P_, B_, flags);
        RH_HIP(hipGetLastError());
    }
}

// One step of the boundary path.  gated: Verlet-list mode (pair kicks over the full list; every
// launch checks the list on the device); else cell-list mode, whose pair kicks the caller has
// just enqueued, at clock offset clock_add.
template <int D>
void Engine<D>::launch_boundary_step(uint32_t clock_add, bool gated) {
    if constexpr (D == 2) {
        const uint32_t nb = blocks_for(n_, TPB);
        if (gated) {
            if (P_.nspecies > 1) k_force_list2<true><<<nb, TPB, 0, stream_>>>(P_, B_);
            else k_force_list2<false><<<nb, TPB, 0, stream_>>>(P_, B_);
            RH_HIP(hipGetLastError());
        }
        enqueue_extras((gated ? EX_GATED : 0u) | EX_PISTON | EX_BOOK);
        if (gated) k_advance2<true><<<nb, TPB, 0, stream_>>>(P_, B_, 0u);
        else k_advance2<false><<<nb, TPB, 0, stream_>>>(P_, B_, clock_add);
        RH_HIP(hipGetLastError());
        if (P_.piston) {
            k_piston2<<<1, 1, 0, stream_>>>(P_, B_, gated ? 1 : 0);
            RH_HIP(hipGetLastError());
        }
    }
}

template <int D>
void Engine<D>::read_compartment_ledger(uint32_t per_wall, double* energy, double* impulse, uint64_t* hits) {
    RH_REQUIRE(per_wall >= 1 && per_wall <= (uint32_t)LEDGER2_COMPARTMENTS, "per_wall must be 1..256");
    std::vector<unsigned long long> L(7 * LEDGER2_COMPARTMENTS * 3, 0ull);
    if (P_.feat && B_.cledger) {
        RH_HIP(hipMemcpyAsync(L.data(), B_.cledger, L.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
        RH_HIP(hipStreamSynchronize(stream_));
    } else {
        // no compartments: every exchange is compartment 0's
        unsigned long long tot[LEDGER_WORDS];
        RH_HIP(hipMemcpyAsync(tot, B_.ledger, sizeof tot, hipMemcpyDeviceToHost, stream_));
        RH_HIP(hipStreamSynchronize(stream_));
        for (int w = 0; w < 7; w++)
            for (int f = 0; f < 3; f++) L[3 * (w * LEDGER2_COMPARTMENTS) + f] = tot[3 * w + f];
    }
    for (uint32_t w = 0; w < 7; w++)
        for (uint32_t c = 0; c < per_wall; c++) {
            const unsigned long long* e = &L[3 * (w * LEDGER2_COMPARTMENTS + c)];
            if (energy) energy[w * per_wall + c] = (double)(long long)e[0] / LEDGER_ONE;
            if (impulse) impulse[w * per_wall + c] = (double)(long long)e[1] / LEDGER_ONE;
            if (hits) hits[w * per_wall + c] = e[2];
        }
}

template <int D>
void Engine<D>::read_piston(double out[4]) {
    RH_REQUIRE(B_.piston != nullptr, "the system has no piston");
    PistonDev p;
    RH_HIP(hipMemcpyAsync(&p, B_.piston, sizeof p, hipMemcpyDeviceToHost, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
    out[0] = core::to_f64(p.w);
    out[1] = (double)p.v * core::UNIT / sys_.dt;
    out[2] = p.force;
    out[3] = p.work;
}

template <int D>
void Engine<D>::set_piston_speed(double u) {
    RH_REQUIRE(B_.piston != nullptr, "the system has no piston");
    const int32_t v = core::to_step_displacement(sys_.dt, u);
    RH_HIP(hipMemcpyAsync(&B_.piston->v, &v, sizeof v, hipMemcpyHostToDevice, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
void Engine<D>::flip_bit(uint32_t particle, uint32_t axis, uint32_t bit) {
    RH_REQUIRE(particle < n_, "particle id out of range");
    flip(particle, 0, axis, bit);
}

template <int D>
void Engine<D>::flip_species_bit(uint32_t species, uint32_t axis, uint32_t bit) { flip(species, 1, axis, bit); }

// Check every flip against the box first, then apply them all (or none), and rebuild the list:
// a flipped particle may have moved any distance.
template <int D>
void Engine<D>::flip(uint32_t sel, int by_species, uint32_t axis, uint32_t bit) {
    if constexpr (D == 2) {
        RH_REQUIRE(axis < 2, "axis must be 0 or 1");
        // Any of the 64 bits, as the reference flips; a flip that leaves the box (every bit from
        // 52 up, for any box the engine accepts) is refused by the check pass.
        RH_REQUIRE(bit < 64, "bit must be below 64");
        const uint32_t nb = blocks_for(n_, TPB);
        write_ctl(CTL_FLIP_BAD, 0u);
        k_flip2<<<nb, TPB, 0, stream_>>>(P_, B_, sel, by_species, axis, bit, 1);
        RH_HIP(hipGetLastError());
        uint32_t ctl[CTL_WORDS];
        read_ctl(ctl);
        RH_REQUIRE(ctl[CTL_FLIP_BAD] == 0u, "the flip would put a particle outside the box");
        k_flip2<<<nb, TPB, 0, stream_>>>(P_, B_, sel, by_species, axis, bit, 0);
        RH_HIP(hipGetLastError());
        write_ctl(CTL_STALE, 1u);
        RH_HIP(hipStreamSynchronize(stream_));
    } else {
        RH_REQUIRE(false, "bit flips are a 2D engine command");
    }
}

// GpuEngine::add_bond / remove_bond (engine.rs:976-1011)
template <int D>
bool Engine<D>::add_bond(uint32_t a, uint32_t b, uint32_t type) {
    RH_REQUIRE(P_.bonds && type < sys_.nbond_types, "add_bond needs a system built with that bond type");
    RH_REQUIRE(a < n_ && b < n_, "particle id out of range");
    auto& bonds = sys_.bonds;
    uint32_t ca = 0, cb = 0;
    for (size_t k = 0; k + 2 < bonds.size(); k += 3) {
        if ((bonds[k] == a && bonds[k + 1] == b) || (bonds[k] == b && bonds[k + 1] == a)) return false;
        ca += (bonds[k] == a) + (bonds[k + 1] == a);
        cb += (bonds[k] == b) + (bonds[k + 1] == b);
    }
    if (a == b || ca >= 4 || cb >= 4) return false;
    bonds.insert(bonds.end(), {a, b, type});
    write_partners();
    return true;
}

template <int D>
bool Engine<D>::remove_bond(uint32_t a, uint32_t b) {
    if (!P_.bonds) return false;
    auto& bonds = sys_.bonds;
    std::vector<uint32_t> keep;
    keep.reserve(bonds.size());
    for (size_t k = 0; k + 2 < bonds.size(); k += 3)
        if (!((bonds[k] == a && bonds[k + 1] == b) || (bonds[k] == b && bonds[k + 1] == a))) keep.insert(keep.end(), {bonds[k], bonds[k + 1], bonds[k + 2]});
    if (keep.size() == bonds.size()) return false;
    bonds.swap(keep);
    write_partners();
    return true;
}

// GpuEngine::sort_now (engine.rs:943-948): reorder storage (and rebuild the list) now.
template <int D>
void Engine<D>::sort_now() {
    enqueue_rebuild(false);
    since_sort_ = 0;
    RH_HIP(hipStreamSynchronize(stream_));
}

template <int D>
void Engine<D>::read_positions(int64_t* pos) { read_state(pos, nullptr, nullptr); }

template class Engine<2>;
template class Engine<3>;

// ---------------------------------------------------------------------------------------------
// Device-side conversion check (used by the tests to pin the kick conversion on the hardware).
// ---------------------------------------------------------------------------------------------

__global__ void k_kick_convert(const float* x, int32_t* out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = cvt_sat_i32(__builtin_rintf(x[i]));
}

// Every f32 r2 in [r2_lo, r2_hi): the shared-reciprocal path of pair_fr against IEEE division,
// bit for bit (fr depends only on r2 for a given species pair, so this is exhaustive).
__global__ void k_check_division(uint32_t lo_bits, uint32_t count, float sigma2, float six_eps4, unsigned long long* mism, uint32_t* first) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        float r2 = __uint_as_float(lo_bits + k);
        float y = rcp_refined(r2);
        float s2f = div_refined(sigma2, r2, y);
        float s6f = s2f * s2f * s2f;
        float frf = div_refined(six_eps4 * s6f * __builtin_fmaf(2.0f, s6f, -1.0f), r2, y);
        float s2 = sigma2 / r2;
        float s6 = s2 * s2 * s2;
        float fr = six_eps4 * s6 * __builtin_fmaf(2.0f, s6, -1.0f) / r2;
        if (__float_as_uint(frf) != __float_as_uint(fr) || __float_as_uint(s2f) != __float_as_uint(s2)) {
            atomicAdd(mism, 1ull);
            atomicMin(first, lo_bits + k);
        }
    }
}

void check_division(float r2_lo, float r2_hi, float sigma2, float six_eps4, uint64_t* checked, uint64_t* mismatches, float* first_bad, hipStream_t s) {
    uint32_t lo = __builtin_bit_cast(uint32_t, r2_lo), hi = __builtin_bit_cast(uint32_t, r2_hi);
    RH_REQUIRE(r2_lo > 0.0f && r2_hi > r2_lo, "need 0 < r2_lo < r2_hi");
    uint32_t count = hi - lo;
    unsigned long long* d_m = nullptr;
    uint32_t* d_f = nullptr;
    RH_HIP(hipMalloc(&d_m, 8));
    RH_HIP(hipMalloc(&d_f, 4));
    RH_HIP(hipMemsetAsync(d_m, 0, 8, s));
    RH_HIP(hipMemsetAsync(d_f, 0xff, 4, s));
    k_check_division<<<4096, 256, 0, s>>>(lo, count, sigma2, six_eps4, d_m, d_f);
    RH_HIP(hipGetLastError());
    unsigned long long m = 0;
    uint32_t f = 0;
    RH_HIP(hipMemcpyAsync(&m, d_m, 8, hipMemcpyDeviceToHost, s));
    RH_HIP(hipMemcpyAsync(&f, d_f, 4, hipMemcpyDeviceToHost, s));
    RH_HIP(hipStreamSynchronize(s));
    (void)hipFree(d_m);
    (void)hipFree(d_f);
    *checked = count;
    *mismatches = m;
    *first_bad = m ? __builtin_bit_cast(float, f) : 0.0f;
}

// The stochastic boundaries' draws on the device, for keys (id0 + k, clock): per key the
// Andersen bath's three velocities and a low heat-bath wall's normal and two tangential ones,
// each as the integer displacement per step the leapfrog stores (the engines' own andersen_draw
// and thermal_wall_draw, with every key redrawn).
struct DrawParams {
    uint32_t seed_lo, seed_hi, bath_species, pinned;
    double bath_p, dt;
    double bath_sd[4];
    double wall_sd[6][4];
};

__global__ void k_stochastic_draws(DrawParams P, uint32_t id0, uint64_t count, uint32_t clock, int32_t* out) {
    const uint64_t k = (uint64_t)blockIdx.x * TPB + threadIdx.x;
    if (k >= count) return;
    const uint32_t id = id0 + (uint32_t)k;
    double u[3] = {0.0, 0.0, 0.0};
    andersen_draw(P, id, clock, 0u, u);
    double vn, t0, t1;
    thermal_wall_draw(P, id, clock, 0, 0, 0u, vn, t0, t1);
    int32_t* o = out + 6 * k;
    o[0] = to_step(P, u[0]);
    o[1] = to_step(P, u[1]);
    o[2] = to_step(P, u[2]);
    o[3] = to_step(P, vn);
    o[4] = to_step(P, t0);
    o[5] = to_step(P, t1);
}

void stochastic_draws(uint64_t seed, double t, double mass, double dt, uint32_t id0, uint64_t count, uint32_t clock, int32_t* out, hipStream_t s) {
    RH_REQUIRE(t > 0.0 && mass > 0.0 && dt > 0.0, "temperature, mass and dt must be positive");
    if (count == 0) return;
    DrawParams P{};
    P.seed_lo = (uint32_t)seed;
    P.seed_hi = (uint32_t)(seed >> 32);
    P.bath_species = 1u;
    P.bath_p = 2.0;  // above every uniform draw: each key is redrawn
    P.dt = dt;
    P.bath_sd[0] = std::sqrt(t / mass);
    P.wall_sd[0][0] = std::sqrt(t / mass);
    k_stochastic_draws<<<blocks_for(count, TPB), TPB, 0, s>>>(P, id0, count, clock, out);
    RH_HIP(hipGetLastError());
}

void kick_convert(const float* x, int32_t* out, uint64_t n, hipStream_t s) {
    if (n == 0) return;
    k_kick_convert<<<blocks_for(n, TPB), TPB, 0, s>>>(x, out, n);
    RH_HIP(hipGetLastError());
}

}  // namespace rh
