// Batched engine (see batch.h).  One iteration, for every active system at once:
//   k_bb_decide          per system: rebuild now? (stale list, steps left, no overflow)
//   k_bb_count           per slot of a rebuilding system: its cell (Morton order inside tiles),
//                        rank in the cell (one atomic per run of equal cells in a wave)
//   k_bb_scan            per system (one work-group): cell counts -> cell ranges, counts cleared
//   k_bb_scatter         per slot: move to its cell-sorted slot in the other buffers
//   then, per dimension on graph branches of their own (different systems):
//   k_bb_row_count, k_bb_row_scan, k_bb_row_scatter
//                        3D systems at least three cells wide: the row-ordered copy
//   k_bb_build_rows      their half list over the rows (one launch per kind)
//   k_bb_build           the others' lists over the cell stencil (2D, small 3D boxes, variant 2
//                        and the full list): in-block pairs at their lower-index end, pairs
//                        across blocks at both ends, as tagged chunks per 64-slot wave
//   k_bb_commit_rebuild  per system of the branch's dimension
//   k_bb_step            per block of a system (one launch per kind): its in-block then its
//                        cross-block chunks dealt evenly over the lanes, then per slot the
//                        leapfrog and half-skin check; systems with a walled axis or the bath
//                        run the booking kind, which books the ledger (and draws the 3D bath and
//                        heat-bath walls)
//   k_bb_commit_step     per system, after the branches join
// Every system takes exactly one step per iteration (its rebuild, when due, runs first in the
// same iteration), so the host enqueues iterations without reading anything back and a chunk of
// iterations can be captured once as a hipGraph and replayed.  Systems never interact, and each
// particle's kick is the same integer sum as in the single-system engine, so every system's
// trajectory is bit-identical to running it alone.
#include "batch.h"

#include <algorithm>
#include <cmath>

#include "md_device.h"

namespace rh {

constexpr uint32_t LIST_UNIT = 4096;  // list-arena allocation unit (entries)
constexpr uint32_t CELL_UNIT = 256;   // cell-arena allocation unit (cells)
constexpr size_t SLOT_BYTES = 48 + 24 + 16 + 24 + 64 + 16 + 32 + 8 + 1;  // per slot: pos, vel, species+ids, ref, aos, rel, row copies, key+rank, wave counts (1/8, rounded up)

// ---------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) { return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v))); }

// The key of cell c: its tile's linear index, then its Morton index inside the tile (bit b of
// the in-tile coordinate on axis a at bit D b + a).  The key is a sum of one term per axis.
template <int D>
__device__ __forceinline__ uint32_t axis_key(const BSys& S, int a, uint32_t v) {
    const uint32_t sh = S.tile_sh, l = v & ((1u << sh) - 1u);
    uint32_t stride = 1u, m = 0u;
    for (int b = 0; b < a; b++) stride *= S.nt[b];
#pragma unroll
    for (uint32_t b = 0; b < 4u; b++) m |= ((l >> b) & 1u) << (D * b + a);
    return (((v >> sh) * stride) << (D * sh)) + m;
}

template <int D>
__device__ __forceinline__ uint32_t lin_cell(const BSys& S, const uint32_t c[D]) {
    uint32_t key = 0u;
#pragma unroll
    for (int a = 0; a < D; a++) key += axis_key<D>(S, a, c[a]);
    return key;
}

// The cell of a key (lin_cell's inverse).
template <int D>
__device__ __forceinline__ void cell_of_key(const BSys& S, uint32_t key, uint32_t c[3]) {
    const uint32_t sh = S.tile_sh, in = key & ((1u << (D * sh)) - 1u);
    uint32_t tile = key >> (D * sh);
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = 0u;
#pragma unroll
    for (int a = 0; a < D; a++) {
        uint32_t l = 0u;
#pragma unroll
        for (uint32_t b = 0; b < 4u; b++) l |= ((in >> (D * b + a)) & 1u) << b;
        c[a] = ((tile % S.nt[a]) << sh) | l;
        tile /= S.nt[a];
    }
}

// Offsets in rb_list of the rebuild lists: one per build kind, then all kinds together.
struct RbOff {
    uint32_t off[RB_ALL + 1];
};

// Per system: rebuild now?  A rebuilding system appends its units to its build kind's list and
// to the list of all kinds (the per-slot rebuild kernels walk those, not the whole arena).
// One wave per system: lane 0 decides and reserves room in the rebuild lists, and the wave
// appends the system's units together.
__global__ __launch_bounds__(64) void k_bb_decide(const BSys* S, BCtl* C, uint32_t nsys, uint32_t* list, uint32_t* count, RbOff ro) {
    const uint32_t s = blockIdx.x;
    if (s >= nsys) return;
    BCtl& c = C[s];
    const BSys& d = S[s];
    // every lane reads the same words, none of which this kernel writes
    const bool go = d.active && c.stale && c.remaining && !c.overflow;
    const uint32_t u0 = d.slot0 / BATCH_WG, nu = d.nslots / BATCH_WG;
    __shared__ uint32_t pq[2];
    if (threadIdx.x == 0u) {
        c.do_rebuild = go ? 1u : 0u;
        if (go) {
            c.maxcount = 0u;
            pq[0] = atomicAdd(&count[d.bkind], nu);
            pq[1] = atomicAdd(&count[RB_ALL], nu);
        }
    }
    if (!go) return;
    __syncthreads();
    const uint32_t p = pq[0], q = pq[1];
    for (uint32_t u = threadIdx.x; u < nu; u += 64u) {
        list[ro.off[d.bkind] + p + u] = u0 + u;
        list[ro.off[RB_ALL] + q + u] = u0 + u;
    }
}

// A slot's cell and its rank in the cell.  The slots arrive in the previous cell order, so a
// wave's lanes mostly hold runs of equal cells: each run takes its ranks with one atomic (made by
// the run's first lane) instead of one per slot contending for the same counter.  Every lane of
// the wave takes part; lanes past the system's end count nothing.
template <int D>
__device__ __forceinline__ void bb_count(BBufs& B, const BSys& S, const BCtl& C, uint32_t slot, bool live) {
    uint32_t key = 0xffffffffu;  // lanes past the end form a run of their own
    if (live) {
        double x[D];
        for (int a = 0; a < D; a++) x[a] = B.pos[C.ppos][a][slot];
        uint32_t c[D];
        cell_of<D>(S, x, c);
        key = lin_cell<D>(S, c);
    }
    const uint32_t lane = __lane_id();
    const uint32_t prev = __shfl_up(key, 1, 64);
    const uint64_t heads = ballot64(lane == 0u || prev != key);
    const uint64_t upto = lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull);  // lanes 0 .. lane
    const uint32_t start = 63u - (uint32_t)__clzll(heads & upto);
    const uint64_t after = heads & ~upto;
    const uint32_t end = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
    uint32_t base = 0u;
    if (live && lane == start) base = atomicAdd(&B.cell_count[S.cell0 + key], end - start);
    base = __shfl(base, (int)start, 64);
    if (!live) return;
    B.key[slot] = key;
    B.rank[slot] = base + (lane - start);
}

// The per-slot rebuild kernels loop over the units of the list they are given (work-group b
// takes its entries b, b + grid, ...), whatever its length this iteration.
__global__ __launch_bounds__(BATCH_WG) void k_bb_count(BBufs B, uint32_t off) {
    const uint32_t len = *(volatile const uint32_t*)&B.rb_count[RB_ALL];
    for (uint32_t b = blockIdx.x; b < len; b += gridDim.x) {
        // the unit and its system are the same for the whole work-group: held in scalar registers,
        // the system's fields are read once into scalar registers too
        const uint32_t u = __builtin_amdgcn_readfirstlane(B.rb_list[off + b]);
        const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[u]);
        const BSys& S = B.sys[s];
        const BCtl& C = B.ctl[s];
        const uint32_t slot = u * BATCH_WG + threadIdx.x;
        const bool live = slot - S.slot0 < S.n;
        if (S.dim == 3) bb_count<3>(B, S, C, slot, live);
        else bb_count<2>(B, S, C, slot, live);
    }
}

template <uint32_t NT = BATCH_WG>
__device__ __forceinline__ uint32_t wg_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t y = __shfl_up(x, d, WAVE);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == WAVE - 1) lds[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < (int)(NT / WAVE); k++) {
        uint32_t s = lds[k];
        pre += (k < (int)w) ? s : 0u;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return pre + x - v;
}

__global__ __launch_bounds__(BATCH_WG) void k_bb_scan(BBufs B, uint32_t nsys) {
    uint32_t s = blockIdx.x;
    if (s >= nsys) return;
    const BSys& S = B.sys[s];
    if (!S.active || !B.ctl[s].do_rebuild) return;
    __shared__ uint32_t lds[BATCH_WG / WAVE];
    uint32_t carry = 0;
    uint32_t* cnt = B.cell_count + S.cell0;
    uint32_t* st = B.cell_start + S.cell0;
    uint32_t* en = B.cell_end + S.cell0;
    for (uint32_t base = 0; base < S.ncells; base += 4 * BATCH_WG) {
        uint32_t c0 = base + 4 * threadIdx.x;
        uint32_t v[4];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = c0 + k < S.ncells ? cnt[c0 + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t ex = carry + wg_exclusive_scan(sum, lds, &total);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (c0 + k < S.ncells) {
                st[c0 + k] = ex;
                en[c0 + k] = ex + v[k];
                cnt[c0 + k] = 0u;
            }
            ex += v[k];
        }
        carry += total;
    }
}

template <int D>
__device__ __forceinline__ void bb_scatter(BBufs& B, const BSys& S, const BCtl& C, uint32_t slot) {
    const uint32_t pp = C.ppos, pm = C.pmeta;
    uint32_t key = B.key[slot];
    uint32_t ns = S.slot0 + B.cell_start[S.cell0 + key] + B.rank[slot];
    uint32_t c[3];
    cell_of_key<D>(S, key, c);
    float r[3] = {0.f, 0.f, 0.f};
    double xs[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double x = B.pos[pp][a][slot];
        xs[a] = x;
        B.pos[pp ^ 1][a][ns] = x;
        B.ref[a][ns] = x;
        B.vel[pm ^ 1][a][ns] = B.vel[pm][a][slot];
        r[a] = (float)(x * 0x1p-32 - (double)c[a] * S.cell_w64[a]);
    }
    B.aos[pp ^ 1][ns] = make_double4(xs[0] * 0.0625, xs[1] * 0.0625, xs[2] * 0.0625, 0.0);  // sixteenths
    B.rel[ns] = make_float4(r[0], r[1], r[2], 0.f);
    B.species[pm ^ 1][ns] = B.species[pm][slot];
    B.ids[pm ^ 1][ns] = B.ids[pm][slot];
}

__global__ __launch_bounds__(BATCH_WG) void k_bb_scatter(BBufs B, uint32_t off) {
    const uint32_t len = *(volatile const uint32_t*)&B.rb_count[RB_ALL];
    for (uint32_t b = blockIdx.x; b < len; b += gridDim.x) {
        // the unit and its system are the same for the whole work-group: held in scalar registers,
        // the system's fields are read once into scalar registers too
        const uint32_t u = __builtin_amdgcn_readfirstlane(B.rb_list[off + b]);
        const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[u]);
        const BSys& S = B.sys[s];
        const BCtl& C = B.ctl[s];
        const uint32_t slot = u * BATCH_WG + threadIdx.x;
        if (slot - S.slot0 >= S.n) continue;
        if (S.dim == 3) bb_scatter<3>(B, S, C, slot);
        else bb_scatter<2>(B, S, C, slot);
    }
}

// The list of local slot i (in the freshly scattered buffers): f32 cell-relative separations
// against each species pair's (rc + skin)(1 + 1e-4); an in-block pair (same block of the
// system's block size) belongs to its lower-index end, pairs across blocks are listed at both
// ends; variant 0 lists every pair at both ends.  As in the single-system engine: candidates
// four at a time in groups of 32 through a distance mask, the candidate and hit loops in step
// across the wave, and full chunks stored densely in the wave's region of 64 cap/4 chunks at
// nbr0 (four indices tagged with the owner's index in its block and entries - 1): in-block
// chunks from the region's start, cross-block chunks from its end down; the wave's two chunk
// counts go to wave_chunks[2 w], wave_chunks[2 w + 1].
template <int D, bool MULTI>
__device__ __forceinline__ void bb_build(BBufs& B, const BSys& S, BCtl& C, uint32_t local) {
    // the system's fields used in the loops, in scalar registers (the same for the whole wave)
    const uint32_t variant = uni(S.variant), block = uni(S.block), slot0 = uni(S.slot0), cell0 = uni(S.cell0);
    const uint32_t periodic = uni(S.periodic), cap = uni(S.cap);
    uint32_t nc[3];
    float cw[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nc[a] = uni(S.nc[a]);
        cw[a] = uni(S.cell_w[a]);
    }
    const uint32_t pm = C.pmeta ^ 1u;
    const uint32_t slot = slot0 + local;
    const uint32_t* sp = B.species[pm] + slot0;
    const float4 me = B.rel[slot];
    const float ri[3] = {me.x, me.y, me.z};
    double xi[D];
#pragma unroll
    for (int a = 0; a < D; a++) xi[a] = B.pos[C.ppos ^ 1u][a][slot];
    const uint32_t si = MULTI ? sp[local] : 0u;
    const float list_r2 = uni(S.pt.list_r2[0]);
    uint32_t c[D];
    cell_of<D>(S, xi, c);
    const uint32_t cap4 = cap / 4u, capw = 64u * cap4;
    const uint32_t own = local & (block - 1u);
    uint4* chunks = reinterpret_cast<uint4*>(B.nbr + S.nbr0) + (size_t)(local >> 6) * capw;
    const uint64_t lt = lanemask_lt();
    // the fills change only where every lane of the wave is present; a store that would meet the
    // other end's chunks flags the overflow instead
    uint32_t fill_in = 0u, fill_out = 0u;
    auto store = [&](auto in_c, bool want, const uint4& q) __attribute__((always_inline)) {
        constexpr bool in = decltype(in_c)::value;
        uint32_t& fill = in ? fill_in : fill_out;
        const uint64_t bal = ballot64(want);
        if (bal == 0ull) return;
        if (want) {
            const uint32_t k = fill + (uint32_t)__popcll(bal & lt), other = in ? fill_out : fill_in;
            if (k + other < capw) chunks[in ? k : capw - 1u - k] = q;
            else atomicOr(&C.overflow, 1u);
        }
        fill += (uint32_t)__popcll(bal);
    };
    // the open in-block and cross-block chunks
    uint4 fa = make_uint4(0u, 0u, 0u, 0u), fb = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t* cs = B.cell_start + cell0;
    const uint32_t* ce = B.cell_end + cell0;
    const float4* rel = B.rel + slot0;
    float wl[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) wl[a] = (nc[a] == 1u && ((periodic >> a) & 1u)) ? cw[a] : 0.f;
    const bool any_wrap = (wl[0] != 0.f) | (wl[1] != 0.f) | (wl[2] != 0.f);
    const uint32_t blo = local & ~(block - 1u);
    uint32_t na = 0, nb = 0;
    // Hit masks wait in a three-entry queue per lane (mask, first candidate) and are visited one
    // hit per step of a wave-wide loop, run only when some lane's queue is full and at the end:
    // the loop's length follows the busiest lane's hits over many cells, not the sum over cells
    // of each cell's busiest lane.
    uint32_t m0 = 0u, m1 = 0u, m2 = 0u, b0 = 0u, b1 = 0u, b2 = 0u, qn = 0u;
    auto visit = [&]() __attribute__((always_inline)) {
        bool full = false, in = false;
        if (qn != 0u) {
            const uint32_t b = __builtin_ctz(m0);
            m0 &= m0 - 1u;
            const uint32_t j = b0 + b;
            in = variant != 0 && j - blo < block;
            // the two counts step by selects (a branch on `in` around them is turned into a
            // per-lane indexed pair in scratch memory)
            if (in) put4(fa, na, j);
            else put4(fb, nb, j);
            na += in ? 1u : 0u;
            nb += in ? 0u : 1u;
            full = ((in ? na : nb) & 3u) == 0u;
            if (m0 == 0u) {
                m0 = m1;
                b0 = b1;
                m1 = m2;
                b1 = b2;
                m2 = 0u;
                qn--;
            }
        }
        // a chunk fills on one step in four: one ballot skips both stores on the others
        if (ballot64(full) != 0ull) {
            const bool fin = full && in, fout = full && !in;
            store(std::true_type{}, fin, tag_chunk(fa, own, 4u, true));
            store(std::false_type{}, fout, tag_chunk(fb, own, 4u, false));
        }
    };
    // each axis's three neighbour coordinates (wrapped on periodic axes) as key terms, and which
    // of them to skip (past a walled face, or off a single-cell axis)
    uint32_t kt[3][3] = {};
    uint32_t skipm = 0u;
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
        for (int o3 = 0; o3 < 3; o3++) {
            int v = (int)c[a] + o3 - 1;
            bool sk = nc[a] == 1u && o3 != 1;
            if (v < 0 || v >= (int)nc[a]) {
                if (!((periodic >> a) & 1u)) sk = true;
                v = v < 0 ? v + (int)nc[a] : v - (int)nc[a];
            }
            kt[a][o3] = axis_key<D>(S, a, (uint32_t)v);
            skipm |= sk ? 1u << (3 * a + o3) : 0u;
        }
    for (int o = 0; o < stencil_size(D); o++) {
        const int o3[3] = {o % 3, o / 3 % 3, o / 9};
        float shift[3] = {0.f, 0.f, 0.f};
        bool skip = false;
        uint32_t key = 0u;
#pragma unroll
        for (int a = 0; a < D; a++) {
            key += o3[a] == 0 ? kt[a][0] : (o3[a] == 1 ? kt[a][1] : kt[a][2]);
            skip |= ((skipm >> (3 * a + o3[a])) & 1u) != 0u;
            shift[a] = (float)(o3[a] - 1) * cw[a] - ri[a];
        }
        uint32_t base = 0u, e0 = 0u;
        if (!skip) {
            const uint32_t s0 = cs[key];
            e0 = ce[key];
            // in-block candidates up to local are not this end's: scan such a cell from past it
            base = (variant != 0 && s0 >= blo) ? max(s0, local + 1u) : s0;
        }
        const float sx = shift[0], sy = shift[1], sz = shift[2];
        for (;; base += 32u) {
            const bool act = base < e0;
            if (ballot64(act) == 0ull) break;
            uint32_t mask = 0u, inb = 0u;
            if (act) {
                const uint32_t cnt = min(32u, e0 - base);
                // four at a time, unrolled; a lane past its cell's end tests the rel entries that
                // follow (the array is padded by 32) and their bits are masked off after
                const float4* __restrict__ rq = rel + base;
                auto near = [&](uint32_t u) -> bool {
                    const float4 q = rq[u];
                    float dx = q.x + sx, dy = q.y + sy, dz = D == 3 ? q.z + sz : 0.f;
                    if (any_wrap) {
                        if (wl[0] != 0.f) dx = dx > 0.5f * wl[0] ? dx - wl[0] : (dx < -0.5f * wl[0] ? dx + wl[0] : dx);
                        if (wl[1] != 0.f) dy = dy > 0.5f * wl[1] ? dy - wl[1] : (dy < -0.5f * wl[1] ? dy + wl[1] : dy);
                        if (D == 3 && wl[2] != 0.f) dz = dz > 0.5f * wl[2] ? dz - wl[2] : (dz < -0.5f * wl[2] ? dz + wl[2] : dz);
                    }
                    float r2 = dx * dx;
                    r2 = __builtin_fmaf(dy, dy, r2);
                    if (D == 3) r2 = __builtin_fmaf(dz, dz, r2);
                    const float lr2 = MULTI ? S.pt.list_r2[si * 4 + sp[min(base + u, e0 - 1u)]] : list_r2;
                    return r2 < lr2;
                };
#pragma unroll 1
                for (uint32_t u0 = 0; u0 < 32u; u0 += 4u) {
                    if (u0 >= cnt) break;
#pragma unroll
                    for (uint32_t k = 0; k < 4u; k++) mask |= near(u0 + k) ? 1u << (u0 + k) : 0u;
                }
                mask &= bits_below((int)cnt);
                // j = base + u: gt: j > local; inb: j in local's block (half list)
                const uint32_t gt = ~bits_below((int)local + 1 - (int)base);
                if (variant != 0) {
                    inb = bits_below((int)(blo + block) - (int)base) & ~bits_below((int)blo - (int)base);
                    mask &= (inb & gt) | ~inb;
                }
                const int t0 = (int)local - (int)base;
                mask &= ~((t0 >= 0 && t0 < 32) ? (1u << t0) : 0u);
            }
            const bool push = mask != 0u;
            while (ballot64(push && qn == 3u) != 0ull) visit();
            if (push) {
                if (qn == 0u) {
                    m0 = mask;
                    b0 = base;
                } else if (qn == 1u) {
                    m1 = mask;
                    b1 = base;
                } else {
                    m2 = mask;
                    b2 = base;
                }
                qn++;
            }
        }
    }
    while (ballot64(qn != 0u) != 0ull) visit();
    const bool pin = (na & 3u) != 0u, pout = (nb & 3u) != 0u;
    store(std::true_type{}, pin, tag_chunk(fa, own, pin ? na & 3u : 1u, true));
    store(std::false_type{}, pout, tag_chunk(fb, own, pout ? nb & 3u : 1u, false));
    const uint32_t ca = (na + 3u) / 4u, cb = (nb + 3u) / 4u;
    if (ca + cb > cap4) atomicOr(&C.overflow, 1u);
    if (4u * (ca + cb) > *(volatile const uint32_t*)&C.maxcount) atomicMax(&C.maxcount, 4u * (ca + cb));
    if (__lane_id() == 0u) {
        const uint32_t in = min(fill_in, capw);
        B.wave_chunks[2u * (slot >> 6)] = in;
        B.wave_chunks[2u * (slot >> 6) + 1u] = min(fill_out, capw - in);
    }
}

// One launch per build kind present (each keeps only its own registers), over that kind's
// rebuild list.
template <int D, bool MULTI>
__global__ __launch_bounds__(BATCH_WG) __attribute__((amdgpu_waves_per_eu(5))) void k_bb_build(BBufs B, uint32_t off) {
    constexpr uint32_t K = (D == 3 ? 2u : 0u) + (MULTI ? 1u : 0u);
    const uint32_t len = *(volatile const uint32_t*)&B.rb_count[K];
    for (uint32_t b = blockIdx.x; b < len; b += gridDim.x) {
        // the unit and its system are the same for the whole work-group: held in scalar registers,
        // the system's fields are read once into scalar registers too
        const uint32_t u = __builtin_amdgcn_readfirstlane(B.rb_list[off + b]);
        const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[u]);
        const BSys& S = B.sys[s];
        BCtl& C = B.ctl[s];
        const uint32_t local = u * BATCH_WG + threadIdx.x - S.slot0;
        if (local >= S.n) {
            // a wave past the system's last particle holds no chunks
            if ((threadIdx.x & 63u) == 0u) {
                const uint32_t w = (u * BATCH_WG + threadIdx.x) >> 6;
                B.wave_chunks[2u * w] = 0u;
                B.wave_chunks[2u * w + 1u] = 0u;
            }
            continue;
        }
        bb_build<D, MULTI>(B, S, C, local);
    }
}

// The systems of dimension `dim` that rebuilt this iteration.
__global__ void k_bb_commit_rebuild(const BSys* S, BCtl* C, uint32_t nsys, uint32_t dim) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys || S[s].dim != dim) return;
    BCtl& c = C[s];
    if (c.do_rebuild) {
        c.stale = 0u;
        c.stale_next = 0u;
        c.ppos ^= 1u;
        c.pmeta ^= 1u;
        c.rebuilds += 1u;
        c.do_rebuild = 0u;
    }
}

// ---------------------------------------------------------------------------------------------
// The 3D list build over the row-ordered copy (systems with rows != 0), as the single-system
// engine's: after the cell sort, each slot's rank in its fine x bin (k_bb_row_count), the
// extended bins' starts per system (k_bb_row_scan, which also clears the counts), the copies
// with their ghosts (k_bb_row_scatter), then the build (k_bb_build_rows).  The slot order stays
// the tiled cell order the step's blocks want; only the build reads the rows.
// ---------------------------------------------------------------------------------------------

// A slot's row (its (y, z) cell) and fine x bin.
__device__ __forceinline__ void bb_row_bin(const BSys& S, const double x[3], uint32_t& row, uint32_t& xf) {
    uint32_t c[3];
    cell_of<3>(S, x, c);
    row = c[2] * S.nc[1] + c[1];
    const double f = x[0] * S.rows_xscale;
    const uint32_t v = f > 0.0 ? (uint32_t)f : 0u;
    xf = v < S.rows_nxf - 1u ? v : S.rows_nxf - 1u;
}

// Over the units of rebuild list K (a rows kind): the slot's rank in its fine bin.
__global__ __launch_bounds__(BATCH_WG) void k_bb_row_count(BBufs B, uint32_t off, uint32_t K) {
    const uint32_t len = *(volatile const uint32_t*)&B.rb_count[K];
    for (uint32_t b = blockIdx.x; b < len; b += gridDim.x) {
        const uint32_t u = __builtin_amdgcn_readfirstlane(B.rb_list[off + b]);
        const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[u]);
        const BSys& S = B.sys[s];
        const BCtl& C = B.ctl[s];
        const uint32_t slot = u * BATCH_WG + threadIdx.x;
        if (slot - S.slot0 >= S.n) continue;
        double x[3];
#pragma unroll
        for (int a = 0; a < 3; a++) x[a] = B.pos[C.ppos ^ 1u][a][slot];
        uint32_t row, xf;
        bb_row_bin(S, x, row, xf);
        B.rank[slot] = atomicAdd(&B.cell_count[S.cell0 + S.ncells + row * S.rows_nxf + xf], 1u);
    }
}

// One work-group of ROW_SCAN_WG lanes per rebuilding rows system: the start of every extended
// bin (a ghost bin takes its source bin's count) and, past the last, the total; then the counts
// are cleared.
constexpr uint32_t ROW_SCAN_WG = 256;
__global__ __launch_bounds__(ROW_SCAN_WG) void k_bb_row_scan(BBufs B, uint32_t nsys) {
    const uint32_t s = blockIdx.x;
    if (s >= nsys) return;
    const BSys& S = B.sys[s];
    if (!S.active || !S.rows || !B.ctl[s].do_rebuild) return;
    __shared__ uint32_t lds[ROW_SCAN_WG / WAVE];
    uint32_t* cnt = B.cell_count + S.cell0 + S.ncells;
    uint32_t* st = B.cell_start + S.cell0 + S.ncells;
    const uint32_t ne = S.rows_ne, nxf = S.rows_nxf, g = S.rows_g;
    const uint32_t nrows = S.nc[1] * S.nc[2], nk = nrows * ne;
    // sixteen consecutive bins per lane: their sixteen count reads are in flight together
    constexpr uint32_t PER = 16;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nk; base += PER * ROW_SCAN_WG) {
        const uint32_t c0 = base + PER * threadIdx.x;
        uint32_t row = c0 / ne, e = c0 - row * ne;
        uint32_t v[PER];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            int b = (int)e - (int)g;  // the source bin of extended bin e (a ghost: the other end's)
            if (b < 0) b += (int)nxf;
            else if (b >= (int)nxf) b -= (int)nxf;
            v[k] = c0 + k < nk ? cnt[row * nxf + (uint32_t)b] : 0u;
            sum += v[k];
            if (++e == ne) {
                e = 0u;
                row++;
            }
        }
        uint32_t total;
        uint32_t ex = carry + wg_exclusive_scan<ROW_SCAN_WG>(sum, lds, &total);
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            if (c0 + k < nk) st[c0 + k] = ex;
            ex += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0u) st[nk] = carry;
    __syncthreads();  // every count is read before any is cleared
    for (uint32_t k = threadIdx.x; k < nrows * nxf; k += ROW_SCAN_WG) cnt[k] = 0u;
}

// The copy of each slot (x absolute in sigma, y and z relative to its cell's corner, its local
// slot with the species in bits 30-31) and its ghost past the row's other end on a periodic x
// axis.
__global__ __launch_bounds__(BATCH_WG) void k_bb_row_scatter(BBufs B, uint32_t off, uint32_t K) {
    const uint32_t len = *(volatile const uint32_t*)&B.rb_count[K];
    for (uint32_t b = blockIdx.x; b < len; b += gridDim.x) {
        const uint32_t u = __builtin_amdgcn_readfirstlane(B.rb_list[off + b]);
        const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[u]);
        const BSys& S = B.sys[s];
        const BCtl& C = B.ctl[s];
        const uint32_t slot = u * BATCH_WG + threadIdx.x;
        const uint32_t local = slot - S.slot0;
        if (local >= S.n) continue;
        double x[3];
#pragma unroll
        for (int a = 0; a < 3; a++) x[a] = B.pos[C.ppos ^ 1u][a][slot];
        uint32_t row, xf;
        bb_row_bin(S, x, row, xf);
        const float4 r = B.rel[slot];
        const uint32_t* rs = B.cell_start + S.cell0 + S.ncells;
        float4* it = B.items + 2 * (size_t)S.slot0;
        const uint32_t rank = B.rank[slot], base = row * S.rows_ne, g = S.rows_g, nxf = S.rows_nxf;
        const uint32_t sp = S.nspecies > 1 ? B.species[C.pmeta ^ 1u][slot] << 30 : 0u;
        const float w = __uint_as_float(local | sp);
        it[rs[base + g + xf] + rank] = make_float4((float)(x[0] * 0x1p-32), r.y, r.z, w);
        if (xf < g)  // a ghost past the row's end
            it[rs[base + g + nxf + xf] + rank] = make_float4((float)((x[0] + S.boxf[0]) * 0x1p-32), r.y, r.z, w);
        if (xf + g >= nxf)  // a ghost before the row's start
            it[rs[base + xf + g - nxf] + rank] = make_float4((float)((x[0] - S.boxf[0]) * 0x1p-32), r.y, r.z, w);
    }
}

// The half list of local slot `local` from the rows: the list sphere meets at most the nine rows
// around the particle's own, and a row's candidates within the sphere's chord are one range of
// the copy (ghost bins make this true across a periodic x face).  Candidates four at a time; the
// three tests (within the radius, in the block, of higher index) are wave masks, shifted into
// per-lane hit masks.  Hit masks wait in a per-lane ring in LDS; a visit moves up to four hits of
// the front mask into that kind's open chunk (also in LDS) and stores chunks as they fill, at
// wave-uniform places of the wave's region: in-block chunks from its start, cross-block chunks
// from its end down, both four indices tagged with the owner (as the stencil build).
template <bool MULTI>
__device__ __forceinline__ void bb_build_rows(BBufs& B, const BSys& S, BCtl& C, uint32_t local) {
    const uint32_t block = uni(S.block), slot0 = uni(S.slot0), cap = uni(S.cap);
    const uint32_t periodic = uni(S.periodic), ny = uni(S.nc[1]), nz = uni(S.nc[2]);
    const uint32_t rows_ne = uni(S.rows_ne), rows_g = uni(S.rows_g), rows_nxf = uni(S.rows_nxf);
    const float wy = uni(S.cell_w[1]), wz = uni(S.cell_w[2]), inv_wx = uni(S.rows_inv_wx), reach = uni(S.rows_reach);
    const uint32_t pm = C.pmeta ^ 1u;
    const uint32_t slot = slot0 + local;
    uint32_t c[3];
    float Xi;
    {
        double xi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) xi[a] = B.pos[C.ppos ^ 1u][a][slot];
        cell_of<3>(S, xi, c);
        Xi = (float)(xi[0] * 0x1p-32);
    }
    float mey, mez;  // y and z relative to the cell corner
    {
        const float4 me = B.rel[slot];
        mey = me.y;
        mez = me.z;
    }
    float lr2[4];
    if constexpr (MULTI) {
        const uint32_t si = B.species[pm][slot];
#pragma unroll
        for (uint32_t s = 0; s < 4; s++) lr2[s] = S.pt.list_r2[si * 4 + s];
    } else {
        lr2[0] = uni(S.pt.list_r2[0]);
    }
    const float reach2 = reach * reach;
    const float4* __restrict__ items = B.items + 2 * (size_t)slot0;
    const uint32_t* __restrict__ rstart = B.cell_start + uni(S.cell0) + uni(S.ncells);
    const uint32_t cap4 = cap / 4u, capw = 64u * cap4;
    const uint32_t blo = local & ~(block - 1u), own = local & (block - 1u);
    uint4* chunks = reinterpret_cast<uint4*>(B.nbr + S.nbr0) + (size_t)(local >> 6) * capw;
    const uint64_t lt = lanemask_lt();
    // the fills change only where every lane of the wave is present; a store that would meet the
    // other end's chunks flags the overflow instead
    uint32_t fill_in = 0u, fill_out = 0u;
    auto store = [&](auto in_c, bool want, const uint4& q) __attribute__((always_inline)) {
        constexpr bool in = decltype(in_c)::value;
        uint32_t& fill = in ? fill_in : fill_out;
        const uint64_t bal = ballot64(want);
        if (bal == 0ull) return;
        if (want) {
            const uint32_t k = fill + (uint32_t)__popcll(bal & lt), other = in ? fill_out : fill_in;
            if (k + other < capw) chunks[in ? k : capw - 1u - k] = q;
            else atomicOr(&C.overflow, 1u);
        }
        fill += (uint32_t)__popcll(bal);
    };
    // Per lane, in LDS: the open in-block and cross-block chunks (entry s at ci[s][t], co[s][t])
    // and a ring of up to four hit masks, each with the copy index of its bit 0 (bit b stands for
    // that index minus b) and its kind in bit 31.
    __shared__ uint32_t ci[4][BATCH_WG], co[4][BATCH_WG];
    __shared__ uint32_t qm[4][BATCH_WG], qt[4][BATCH_WG];
    const uint32_t t = threadIdx.x;
    uint32_t na = 0u, nb = 0u, head = 0u, qn = 0u;
    auto chunk_of = [&](bool in, uint32_t entries) {
        return in ? tag_chunk(make_uint4(ci[0][t], ci[1][t], ci[2][t], ci[3][t]), own, entries, true)
                  : tag_chunk(make_uint4(co[0][t], co[1][t], co[2][t], co[3][t]), own, entries, false);
    };
    auto visit = [&]() __attribute__((always_inline)) {
        bool full = false, in = false;
        if (qn != 0u) {
            const uint32_t h = head & 3u;
            uint32_t m = qm[h][t];
            const uint32_t tag = qt[h][t];
            in = (tag >> 31) != 0u;
            const uint32_t b1 = tag & 0x7fffffffu;  // the copy index of bit 0
            const uint32_t n = in ? na : nb, r = n & 3u;
            uint32_t jv[4];
            bool take[4];
#pragma unroll
            for (uint32_t s2 = 0; s2 < 4u; s2++) {
                take[s2] = r + s2 < 4u && m != 0u;
                const uint32_t at = take[s2] ? (uint32_t)__builtin_ctz(m) : 0u;
                m = take[s2] ? m & (m - 1u) : m;
                jv[s2] = __float_as_uint(items[b1 - at].w) & 0x0fffffffu;
            }
#pragma unroll
            for (uint32_t s2 = 0; s2 < 4u; s2++) {
                if (!take[s2]) continue;
                if (in) ci[r + s2][t] = jv[s2];
                else co[r + s2][t] = jv[s2];
            }
            const uint32_t taken = (uint32_t)take[0] + (uint32_t)take[1] + (uint32_t)take[2] + (uint32_t)take[3];
            const uint32_t n2 = n + taken;
            full = (n2 & 3u) == 0u;  // at least one was taken, so the chunk closed
            na = in ? n2 : na;
            nb = in ? nb : n2;
            if (m == 0u) {
                head++;
                qn--;
            } else {
                qm[h][t] = m;
            }
        }
        // every lane takes part in the stores' ballots, so the fills stay the same across the wave
        if (ballot64(full) != 0ull) {
            const bool fin = full && in, fout = full && !in;
            store(std::true_type{}, fin, fin ? chunk_of(true, 4u) : make_uint4(0u, 0u, 0u, 0u));
            store(std::false_type{}, fout, fout ? chunk_of(false, 4u) : make_uint4(0u, 0u, 0u, 0u));
        }
    };
    auto push = [&](uint32_t m, uint32_t tag) {
        if (m == 0u) return;
        const uint32_t h = (head + qn) & 3u;
        qm[h][t] = m;
        qt[h][t] = tag;
        qn++;
    };
    // the chord's bins are clamped to the row's extended bins (its ghosts on a periodic x axis)
    const int glo = -(int)rows_g, ghi = (int)(rows_nxf + rows_g) - 1;
    const float sx = -Xi;
    for (int o = 0; o < 9; o++) {
        const int oy = o % 3 - 1, oz = o / 3 - 1;
        int yy = (int)c[1] + oy, zz = (int)c[2] + oz;
        bool skip = false;
        if (yy < 0 || yy >= (int)ny) {
            if (!((periodic >> 1) & 1u)) skip = true;
            yy = yy < 0 ? yy + (int)ny : yy - (int)ny;
        }
        if (zz < 0 || zz >= (int)nz) {
            if (!((periodic >> 2) & 1u)) skip = true;
            zz = zz < 0 ? zz + (int)nz : zz - (int)nz;
        }
        // distance from the particle to the row's (y, z) rectangle, and the chord along x
        const float dyl = oy == 0 ? 0.f : fmaxf(0.f, oy > 0 ? wy - mey : mey);
        const float dzl = oz == 0 ? 0.f : fmaxf(0.f, oz > 0 ? wz - mez : mez);
        const float d2 = dyl * dyl + dzl * dzl;
        skip |= !(d2 < reach2);
        const float ext = skip ? 0.f : __builtin_sqrtf(reach2 - d2);
        const float sy = (float)oy * wy - mey, sz = (float)oz * wz - mez;
        int a = (int)__builtin_floorf((Xi - ext) * inv_wx);
        int b = (int)__builtin_floorf((Xi + ext) * inv_wx);
        a = a > glo ? a : glo;
        b = b < ghi ? b : ghi;
        uint32_t s0 = 0u, e0 = 0u;
        if (!skip && a <= b) {
            const uint32_t rowbase = ((uint32_t)zz * ny + (uint32_t)yy) * rows_ne + rows_g;
            s0 = rstart[rowbase + (uint32_t)a];
            e0 = rstart[rowbase + (uint32_t)b + 1u];
        }
        // Candidates in groups of up to 32, tested without a branch (each candidate is one
        // 16-byte load) into two masks: in-block partners of higher index than the particle (an
        // in-block pair is kept at its lower-index end only) and partners outside its block,
        // both within the list radius.
        for (uint32_t base = s0;; base += 32u) {
            const bool act = base < e0;
            if (ballot64(act) == 0ull) break;
            // Hits are shifted in as they are tested: candidate base + u ends at bit n - 1 - u of
            // its mask, n being the number of candidates tested.
            uint32_t m_in = 0u, m_out = 0u, last = base;
            if (act) {
                const uint32_t cnt = min(32u, e0 - base);
                // four at a time; a lane past its range's end tests the copy's entries that follow
                // (the copy is padded by 32), whose bits are masked off after
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
                        const uint64_t near = ballot64(r2 < lr), inb = ballot64(j - blo < block), gt = ballot64(j > local);
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
    while (ballot64(qn != 0u) != 0ull) visit();
    // the partial last chunks; a list longer than the capacity is flagged and regrown
    const bool pin = (na & 3u) != 0u, pout = (nb & 3u) != 0u;
    store(std::true_type{}, pin, pin ? chunk_of(true, na & 3u) : make_uint4(0u, 0u, 0u, 0u));
    store(std::false_type{}, pout, pout ? chunk_of(false, nb & 3u) : make_uint4(0u, 0u, 0u, 0u));
    const uint32_t ca = (na + 3u) / 4u, cb = (nb + 3u) / 4u;
    if (ca + cb > cap4) atomicOr(&C.overflow, 1u);
    if (4u * (ca + cb) > *(volatile const uint32_t*)&C.maxcount) atomicMax(&C.maxcount, 4u * (ca + cb));
    if (__lane_id() == 0u) {
        const uint32_t in = min(fill_in, capw);
        B.wave_chunks[2u * (slot >> 6)] = in;
        B.wave_chunks[2u * (slot >> 6) + 1u] = min(fill_out, capw - in);
    }
}

template <bool MULTI>
__global__ __launch_bounds__(BATCH_WG) void k_bb_build_rows(BBufs B, uint32_t off) {
    constexpr uint32_t K = 4u + (MULTI ? 1u : 0u);
    const uint32_t len = *(volatile const uint32_t*)&B.rb_count[K];
    for (uint32_t b = blockIdx.x; b < len; b += gridDim.x) {
        const uint32_t u = __builtin_amdgcn_readfirstlane(B.rb_list[off + b]);
        const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[u]);
        const BSys& S = B.sys[s];
        BCtl& C = B.ctl[s];
        const uint32_t local = u * BATCH_WG + threadIdx.x - S.slot0;
        if (local >= S.n) {
            // a wave past the system's last particle holds no chunks
            if ((threadIdx.x & 63u) == 0u) {
                const uint32_t w = (u * BATCH_WG + threadIdx.x) >> 6;
                B.wave_chunks[2u * w] = 0u;
                B.wave_chunks[2u * w + 1u] = 0u;
            }
            continue;
        }
        bb_build_rows<MULTI>(B, S, C, local);
    }
}

__device__ __forceinline__ bool bb_runs(const BSys& S, const BCtl& C) { return S.active && (C.stale | C.overflow) == 0u && C.remaining != 0u; }

// One work-group of BS slots, a block of the system: as the single-system step kernel, the
// block's chunks of each kind (its waves' regions, numbered in order) are dealt out evenly over
// the lanes, in-block chunks first (partners read from LDS), then cross-block ones (partners
// gathered); every kick goes to the block's LDS accumulators, and positions are staged and
// gathered as sixteenths.
template <int D, bool MULTI, bool BOOK, uint32_t BS>
__device__ __forceinline__ void bb_step(BBufs& B, const BSys& S, BCtl& C, uint32_t sysid, uint32_t wg) {
    constexpr uint32_t NW = BS / 64u;
    __shared__ uint32_t acc[D][BS];
    __shared__ double lpos[D][BS];  // the block's positions in sixteenths, for in-block neighbours
    __shared__ uint32_t wfirst[2][NW + 1];
    __shared__ uint32_t edge;
    const uint32_t t = threadIdx.x;
    const uint32_t slot = wg * BS + t;
    const uint32_t local = slot - S.slot0;
    const uint32_t base = wg * BS - S.slot0;  // first local slot of this block
    const bool live = local < S.n;
    const uint32_t pp = C.ppos, pm = C.pmeta;
    const double* px[D];
    for (int a = 0; a < D; a++) px[a] = B.pos[pp][a] + S.slot0;
    if (t == 0u) edge = S.range_check;
    // Periodic axes on which the block needs no image: with every particle of the block at least
    // face16 from both faces (the bound on a listed partner's distance at step time: the largest
    // list radius plus the two particles' moves since the build), every partner's stored position
    // is its nearest image and separation16's correction is exactly zero.
    uint32_t near_face = 0u;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double x16 = live ? px[a][local] * 0.0625 : 0.0;
        acc[a][t] = 0u;
        lpos[a][t] = x16;
        if (live && ((S.periodic >> a) & 1u) && !(x16 >= S.face16 && x16 <= S.box16[a] - S.face16)) near_face = 1u;
    }
    if (t < 64u) {
#pragma unroll
        for (uint32_t kind = 0; kind < 2u; kind++) {
            const uint32_t v = t < NW ? B.wave_chunks[2u * ((wg * BS >> 6) + t) + kind] : 0u;
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
    if (ballot64(near_face != 0u) != 0ull && __lane_id() == 0u) atomicOr(&edge, 1u);
    __syncthreads();
    const bool images = __builtin_amdgcn_readfirstlane(edge) != 0u;
    const uint32_t* sp = B.species[pm] + S.slot0;
    const PairTable& pt = S.pt;
    {
        const float rc2_0 = pt.rc2[0], sigma2_0 = pt.sigma2[0], six_eps4_0 = pt.six_eps4[0], r2f_0 = pt.r2_fast[0];
        const size_t capw = 64u * (S.cap / 4u);
        const uint4* __restrict__ chunks = reinterpret_cast<const uint4*>(B.nbr + S.nbr0) + (size_t)(base >> 6) * capw;
        const double4* __restrict__ pq = B.aos[pp] + S.slot0;
        const bool rc = S.range_check != 0;
        // one pair: its separation (with or without the periodic image), its kick on the owner
        // (summed into k) and its kick on the partner (kj); false beyond the cutoff or, on a
        // system whose list separations may reach 8 sigma, outside the reference's range
        auto pair = [&](auto with_images, const double xi[D], const double xs[D], uint32_t si, uint32_t j, float c, uint32_t k[D], int32_t kj[D]) -> bool {
            float f[D];
            double d[D];
#pragma unroll
            for (int a = 0; a < D; a++) {
                d[a] = decltype(with_images)::value ? separation16<D>(S, xi[a], xs[a], a) : xs[a] - xi[a];
                f[a] = sep_f16(d[a]);
            }
            if constexpr (decltype(with_images)::value) {
                if (rc) {
                    // the reference's range test on the separation (|d| < 8 sigma, in sixteenths)
                    bool ok = true;
#pragma unroll
                    for (int a = 0; a < D; a++) ok &= d[a] >= -2147483648.0 && d[a] < 2147483648.0;
                    if (!ok) return false;
                }
            }
            if constexpr (MULTI) {
                const uint32_t sj = sp[j], p = si * 4 + sj;
                return pair_kick2<D, true>(f, pt.rc2[p], pt.sigma2[p], pt.six_eps4[p], pt.r2_fast[p], c, pt.kick_c[sj], k, kj);
            } else {
                return pair_kick2<D, false>(f, rc2_0, sigma2_0, six_eps4_0, r2f_0, c, c, k, kj);
            }
        };
        auto own_of = [&](uint32_t own, uint32_t& si, float& c) {
            si = 0u;
            c = S.kick_c[0];
            if (MULTI) {
                si = sp[base + own];
                c = pt.kick_c[si];
            }
        };
        // In-block chunks: owner and partners both in LDS, every kick to the LDS accumulators.
        auto inblock_pass = [&](auto with_images) {
            const uint32_t nch = wfirst[0][NW];
            uint32_t g = t, w = 0u;
            auto at = [&](uint32_t gg) -> size_t {
                while (gg >= wfirst[0][w + 1u]) w++;
                return (size_t)w * capw + (gg - wfirst[0][w]);
            };
            // the next chunk is in flight while this one is evaluated
            uint4 q_next = make_uint4(0u, 0u, 0u, 0u);
            if (g < nch) q_next = chunks[at(g)];
            for (; g < nch; g += BS) {
                const uint4 qc = q_next;
                if (g + BS < nch) q_next = chunks[at(g + BS)];
                uint32_t own, m;
                bool inblock;
                const uint4 q = untag_chunk(qc, own, m, inblock);
                const uint32_t e[4] = {q.x - base, q.y - base, q.z - base, q.w - base};
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
                for (int s2 = 0; s2 < 4; s2++) {
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
        // Cross-block chunks (and every chunk of a full list): partners gathered from the packed
        // positions, owner kicks only (each such pair is listed at both its ends).
        auto cross_pass = [&](auto with_images) {
            const uint32_t nch = wfirst[1][NW];
            uint32_t g = t, w = 0u;
            auto at = [&](uint32_t gg) -> size_t {
                while (gg >= wfirst[1][w + 1u]) w++;
                return (size_t)w * capw + capw - 1u - (gg - wfirst[1][w]);
            };
            uint4 q_next = make_uint4(0u, 0u, 0u, 0u);
            if (g < nch) q_next = chunks[at(g)];
            for (; g < nch; g += BS) {
                const uint4 qc = q_next;
                if (g + BS < nch) q_next = chunks[at(g + BS)];
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
                for (int e = 0; e < 4; e++) {
                    if constexpr (D == 3) {
                        const double4 q4 = pq[js[e]];
                        xs[e][0] = q4.x;
                        xs[e][1] = q4.y;
                        xs[e][2] = q4.z;
                    } else {
                        const double2 q2 = reinterpret_cast<const double2*>(pq + js[e])[0];
                        xs[e][0] = q2.x;
                        xs[e][1] = q2.y;
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if ((uint32_t)e >= m) break;
                    int32_t kj[D];
                    (void)pair(with_images, xi, xs[e], si, js[e], c, k, kj);
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
    if (!live) return;
    uint32_t si = 0;
    float c = S.kick_c[0];
    if (MULTI) {
        si = sp[local];
        c = pt.kick_c[si];
    }
    double xi[D];
    uint32_t k[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        xi[a] = lpos[a][t] * 16.0;
        k[a] = 0u;
    }
    float wf[2 * D] = {};  // each soft wall's force on this particle (booked as impulse)
    if (S.soft_walls) soft_walls<D>(S, xi, c, k, nullptr, BOOK ? wf : nullptr);
#pragma unroll
    for (int a = 0; a < D; a++) k[a] += acc[a][t];
    const bool pin = (S.pinned >> si) & 1u;
    int32_t v[D];
    double x[D];
#pragma unroll
    for (int a = 0; a < D; a++) v[a] = (int32_t)((uint32_t)B.vel[pm][a][slot] + k[a] + (pin ? 0u : (uint32_t)S.gravity[a]));
    if constexpr (BOOK) {
        // as the single-system engine's booked leapfrog: the soft walls' impulses, the Andersen
        // bath (3D) or the thermostat region (2D), then the move and the walls (heat-bath ones
        // drawing the velocity), every
        // exchange booked in the system's ledger (the step's clock is the system's before this
        // step is committed)
        unsigned long long* L = B.ledger + (size_t)sysid * LEDGER_WORDS;
        const uint32_t id = B.ids[pm][slot], clock = C.clock;
#pragma unroll
        for (int w = 0; w < 2 * D; w++)
            if (wf[w] != 0.0f) ledger_add(L, w, 1, (double)wf[w] * S.dt);
        if (S.bath) {
            if constexpr (D == 3) bath_redraw<3>(S, L, id, clock, si, v);
            else thermostat_redraw2(S, L, id, clock, si, xi, v);
        }
#pragma unroll
        for (int a = 0; a < D; a++) x[a] = xi[a] + (double)v[a];
        bound_booked<D>(S, L, x, v, id, clock, si);
    } else {
#pragma unroll
        for (int a = 0; a < D; a++) x[a] = xi[a] + (double)v[a];
        bound<D>(S, x, v);
    }
    float r2 = 0.0f;
    bool far = false;
#pragma unroll
    for (int a = 0; a < D; a++) {
        B.pos[pp ^ 1][a][slot] = x[a];
        B.vel[pm][a][slot] = v[a];
        float d;
        far |= !sep_f(separation<D>(S, B.ref[a][slot], x[a], a), d);
        r2 += d * d;
    }
    B.aos[pp ^ 1][slot] = make_double4(x[0] * 0.0625, x[1] * 0.0625, D == 3 ? x[D - 1] * 0.0625 : 0.0, 0.0);
    bool stale = far || r2 >= S.half_skin2;
    uint64_t bal = ballot64(stale);
    if (bal != 0ull && __lane_id() == (uint32_t)(__ffsll((long long)bal) - 1) && *(volatile const uint32_t*)&C.stale_next == 0u)
        atomicOr(&C.stale_next, 1u);
}

// One launch per kind of system present (dimension, one or more species, booking the ledger or
// not) and block size, each compiled with only its own path so the common kernels keep their
// registers low.
// Work-group k takes block step_list[off + k] of the launch's kind: slots [g BS, (g + 1) BS) of
// a system of block size BS (its slots start at a multiple of its block size and fill whole
// blocks).  Consecutive blocks of the list go to the same XCD.
template <int D, bool MULTI, bool BOOK, uint32_t BS>
__global__ __launch_bounds__(BS) void k_bb_step(BBufs B, uint32_t off) {
    const uint32_t wg = __builtin_amdgcn_readfirstlane(B.step_list[off + xcd_block(blockIdx.x, gridDim.x)]);
    const uint32_t s = __builtin_amdgcn_readfirstlane(B.wg_sys[wg * (BS / BATCH_WG)]);
    const BSys& S = B.sys[s];
    BCtl& C = B.ctl[s];
    if (!bb_runs(S, C)) return;
    bb_step<D, MULTI, BOOK, BS>(B, S, C, s, wg);
}

__global__ void k_bb_commit_step(const BSys* S, BCtl* C, uint32_t nsys, uint32_t* rb_count) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= RB_ALL) rb_count[s] = 0u;  // the next iteration's rebuild lists start empty
    if (s >= nsys) return;
    BCtl& c = C[s];
    if (bb_runs(S[s], c)) {
        c.ppos ^= 1u;
        c.remaining -= 1u;
        uint32_t lo = c.done_lo + 1u;
        c.done_lo = lo;
        if (lo == 0u) c.done_hi += 1u;
        c.clock += 1u;
        c.stale = c.stale_next;
        c.stale_next = 0u;
    } else if (S[s].active && c.remaining) {
        c.skipped += 1u;
    }
}

// ---------------------------------------------------------------------------------------------
// RangeAlloc
// ---------------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------------
// Batch host side
// ---------------------------------------------------------------------------------------------

void* Batch::dalloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) throw Error(std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    allocs_.push_back(p);
    return p;
}

Batch::Batch(uint64_t budget, uint32_t max_systems) : max_sys_(max_systems) {
    RH_REQUIRE(max_systems >= 1 && max_systems <= (1u << 20), "max_systems must be 1..2^20");
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        RH_HIP(hipMemGetInfo(&fr, &tot));
        budget = fr / 2;
    }
    uint64_t slot_cap = (uint64_t)(budget * 0.16 / SLOT_BYTES) / BATCH_MAX_BLOCK * BATCH_MAX_BLOCK;
    uint64_t list_cap = (uint64_t)(budget * 0.80 / 4) / LIST_UNIT * LIST_UNIT;
    uint64_t cell_cap = (uint64_t)(budget * 0.04 / 12) / CELL_UNIT * CELL_UNIT;
    slot_cap = std::min<uint64_t>(slot_cap, (uint64_t(1) << 31) - BATCH_MAX_BLOCK);
    RH_REQUIRE(slot_cap >= BATCH_MAX_BLOCK && list_cap >= LIST_UNIT && cell_cap >= CELL_UNIT, "batch memory budget too small");
    slots_.reset(slot_cap / BATCH_WG);
    lists_.reset(list_cap / LIST_UNIT);
    cells_.reset(cell_cap / CELL_UNIT);
    RH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    RH_HIP(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
    RH_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    RH_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    RH_HIP(hipEventCreateWithFlags(&ev_join2_, hipEventDisableTiming));
    std::memset(&B_, 0, sizeof B_);
    for (int p = 0; p < 2; p++) {
        for (int a = 0; a < 3; a++) {
            B_.pos[p][a] = (double*)dalloc(slot_cap * 8);
            B_.vel[p][a] = (int32_t*)dalloc(slot_cap * 4);
        }
        B_.species[p] = (uint32_t*)dalloc(slot_cap * 4);
        B_.ids[p] = (uint32_t*)dalloc(slot_cap * 4);
        B_.aos[p] = (double4*)dalloc(slot_cap * 32);
    }
    for (int a = 0; a < 3; a++) B_.ref[a] = (double*)dalloc(slot_cap * 8);
    B_.rel = (float4*)dalloc((slot_cap + 32) * 16);  // padded: the list build reads up to 31 past a cell
    B_.items = (float4*)dalloc((2 * slot_cap + 32) * 16);  // two per slot; padded as rel
    RH_HIP(hipMemset(B_.items, 0, (2 * slot_cap + 32) * 16));
    B_.wave_chunks = (uint32_t*)dalloc(slot_cap / 64 * 8);
    RH_HIP(hipMemset(B_.wave_chunks, 0, slot_cap / 64 * 8));
    B_.key = (uint32_t*)dalloc(slot_cap * 4);
    B_.rank = (uint32_t*)dalloc(slot_cap * 4);
    B_.nbr = (uint32_t*)dalloc(list_cap * 4);
    B_.cell_count = (uint32_t*)dalloc(cell_cap * 4);
    B_.cell_start = (uint32_t*)dalloc(cell_cap * 4);
    B_.cell_end = (uint32_t*)dalloc(cell_cap * 4);
    RH_HIP(hipMemset(B_.cell_count, 0, cell_cap * 4));
    d_sys_ = (BSys*)dalloc(sizeof(BSys) * max_sys_);
    d_ctl_ = (BCtl*)dalloc(sizeof(BCtl) * max_sys_);
    RH_HIP(hipMemset(d_sys_, 0, sizeof(BSys) * max_sys_));
    RH_HIP(hipMemset(d_ctl_, 0, sizeof(BCtl) * max_sys_));
    B_.ledger = (unsigned long long*)dalloc(sizeof(unsigned long long) * LEDGER_WORDS * max_sys_);
    RH_HIP(hipMemset(B_.ledger, 0, sizeof(unsigned long long) * LEDGER_WORDS * max_sys_));
    d_wg_ = (uint32_t*)dalloc(sizeof(uint32_t) * (slot_cap / BATCH_WG));
    RH_HIP(hipMemset(d_wg_, 0xff, sizeof(uint32_t) * (slot_cap / BATCH_WG)));
    B_.wg_sys = d_wg_;
    units_cap_ = (uint32_t)(slot_cap / BATCH_WG);
    d_step_list_ = (uint32_t*)dalloc(sizeof(uint32_t) * units_cap_);
    d_rb_list_ = (uint32_t*)dalloc(sizeof(uint32_t) * 2 * (size_t)units_cap_);
    d_rb_count_ = (uint32_t*)dalloc(sizeof(uint32_t) * (RB_ALL + 1));
    RH_HIP(hipMemset(d_rb_count_, 0, sizeof(uint32_t) * (RB_ALL + 1)));
    B_.step_list = d_step_list_;
    B_.rb_list = d_rb_list_;
    B_.rb_count = d_rb_count_;
    B_.sys = d_sys_;
    B_.ctl = d_ctl_;
    h_sys_.assign(max_sys_, BSys{});
    for (int i = (int)max_sys_ - 1; i >= 0; i--) free_ids_.push_back(i);
    RH_HIP(hipHostMalloc((void**)&h_ctl_, sizeof(BCtl) * max_sys_, hipHostMallocDefault));
    RH_HIP(hipDeviceSynchronize());
}

Batch::~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (void* p : allocs_) (void)hipFree(p);
    if (h_ctl_) (void)hipHostFree(h_ctl_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (ev_join2_) (void)hipEventDestroy(ev_join2_);
    if (aux_) (void)hipStreamDestroy(aux_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::budget(uint64_t* slot_cap, uint64_t* list_cap, uint64_t* cell_cap) const {
    if (slot_cap) *slot_cap = slots_.capacity() * BATCH_WG;
    if (list_cap) *list_cap = lists_.capacity() * LIST_UNIT;
    if (cell_cap) *cell_cap = cells_.capacity() * CELL_UNIT;
}

// Cells, list capacity and constants of one system, as Params::new (engine3.rs:70-162).
void Batch::plan(const SystemDesc& sys, uint64_t n, double skin, BSys& d, uint32_t* cap) const {
    RH_REQUIRE(!sys.compartments[0] && !sys.compartments[1] && !sys.has_membrane && !sys.has_piston && sys.bonds.empty(),
               "compartments, membranes, pistons and bonds are not supported by the batched engine");
    const int D = sys.dim;
    RH_REQUIRE(D == 2 || D == 3, "dimension must be 2 or 3");
    RH_REQUIRE(sys.nspecies >= 1 && sys.nspecies <= 4, "between 1 and 4 species");
    // list chunks carry their tag in the top four bits of each index (tag_chunk)
    RH_REQUIRE(n >= 1 && n < (1u << 28), "particle count out of range (below 2^28 per system)");
    RH_REQUIRE(skin >= 0.0, "skin must be >= 0");
    std::memset(&d, 0, sizeof d);
    d.dim = D;
    d.n = (uint32_t)n;
    d.nspecies = sys.nspecies;
    d.pinned = sys.pinned;
    double rcm = core::rc_max(sys.pairs, sys.nspecies);
    RH_REQUIRE(rcm < 7.9, "cutoff must be below 8 sigma");
    double b[3] = {0, 0, 0}, bmax = 0, vol = 1;
    for (int a = 0; a < D; a++) {
        b[a] = core::to_f64(sys.box_len[a]);
        bmax = std::max(bmax, b[a]);
        vol *= b[a];
    }
    double margin = 1e-3 + 1e-6 * bmax;
    uint32_t size_class = 0;
    while (size_class < 31 && (uint64_t(2) << size_class) <= n) size_class++;
    const Tuning& tune = tuning_[D - 2][size_class];
    d.variant = tune.variant;
    d.block = tune.block;
    d.bkind = (D == 3 ? 2u : 0u) + (sys.nspecies > 1 ? 1u : 0u);
    uint64_t ncells = 1;
    for (int a = 0; a < D; a++) {
        d.nc[a] = core::sat_u32(std::floor(b[a] / ((rcm + skin) * tune.cell_factor + margin)));
        if (d.nc[a] < 3) {
            RH_REQUIRE(b[a] > 2.0 * (rcm + skin) * (1.0 + 1e-4) + margin, "box too small: shorter than twice the cutoff plus skin");
            d.nc[a] = 1;
        }
        ncells *= d.nc[a];
        RH_REQUIRE(sys.box_len[a] < (int64_t(1) << 52), "box longer than 2^20 sigma");
        d.box[a] = sys.box_len[a];
        d.half[a] = sys.box_len[a] / 2;
        d.boxf[a] = (double)sys.box_len[a];
        d.inv_box[a] = sys.walled[a] ? 0.0 : 1.0 / d.boxf[a];  // 0 on walled axes: no image
        d.box16[a] = d.boxf[a] * 0.0625;  // the same image on positions held in sixteenths
        d.inv_box16[a] = d.inv_box[a] * 16.0;
        d.cell_scale[a] = (double)d.nc[a] / (double)sys.box_len[a];
        d.cell_w64[a] = b[a] / (double)d.nc[a];
        d.cell_w[a] = (float)d.cell_w64[a];
        if (!sys.walled[a]) d.periodic |= 1u << a;
        else
            for (int side = 0; side < 2; side++) {
                if (sys.wall_kind[a][side] == RH_WALL_SOFT) d.soft_walls |= 1u << (2 * a + side);
                if (sys.wall_kind[a][side] == RH_WALL_THERMAL) {
                    d.thermal_walls |= 1u << (2 * a + side);
                    for (uint32_t sp = 0; sp < sys.nspecies; sp++) d.wall_sd[2 * a + side][sp] = std::sqrt(sys.wall_t[a][side] / sys.masses[sp]);
                }
            }
        d.gravity[a] = core::gravity_kick(sys.dt, sys.gravity[a]);
    }
    // the ledger, as the single-system engine books it (a walled axis or the bath), and the
    // stochastic boundaries' Threefry keys (the system seed)
    d.ledger = (sys.has_bath || d.periodic != (1u << D) - 1u) ? 1u : 0u;
    for (uint32_t sp = 0; sp < sys.nspecies; sp++) d.mass[sp] = sys.masses[sp];
    d.seed_lo = (uint32_t)sys.seed;
    d.seed_hi = (uint32_t)(sys.seed >> 32);
    d.dt = sys.dt;
    if (sys.has_bath) {
        d.bath = 1u;
        d.bath_species = sys.bath_species;
        d.bath_p = sys.bath_rate * sys.dt;
        for (uint32_t sp = 0; sp < sys.nspecies; sp++) d.bath_sd[sp] = std::sqrt(sys.bath_t / sys.masses[sp]);
        if (D == 2) {
            d.bath_axis = sys.bath_axis;
            d.bath_drift = core::to_step_displacement(sys.dt, sys.bath_drift);
            d.bath_lo = (double)sys.bath_region[0];
            d.bath_hi = (double)sys.bath_region[1];
        }
    }
    // Tiles: the smallest power-of-two edge whose cells hold about two blocks of particles (at
    // most 8 cells in 3D, 16 in 2D), halved while padding the grid to whole tiles would more
    // than double the cell count.
    {
        const double per_cell = (double)n / (double)ncells;
        const uint32_t tmax = D == 3 ? 3u : 4u;
        uint32_t sh = 0;
        while (sh < tmax && std::pow((double)(1u << sh), D) * per_cell < 2.0 * d.block) sh++;
        for (;; sh--) {
            uint64_t padded = 1;
            for (int a = 0; a < D; a++) padded *= (uint64_t)((d.nc[a] + (1u << sh) - 1u) >> sh) << sh;
            if (sh == 0 || padded <= 2 * ncells) {
                ncells = padded;
                break;
            }
        }
        d.tile_sh = sh;
        for (int a = 0; a < 3; a++) d.nt[a] = a < D ? (d.nc[a] + (1u << sh) - 1u) >> sh : 1u;
    }
    RH_REQUIRE(ncells < (uint64_t(1) << 31), "too many cells");
    d.ncells = (uint32_t)ncells;
    for (uint32_t s = 0; s < 4; s++) {
        float kc = s < sys.nspecies ? (float)core::kick_coeff(sys.dt, sys.masses[s], (sys.pinned >> s) & 1u) : 0.0f;
        d.kick_c[s] = kc;
        d.pt.kick_c[s] = kc;
    }
    for (uint32_t a = 0; a < sys.nspecies; a++)
        for (uint32_t bb = 0; bb < sys.nspecies; bb++) {
            int k = a * 4 + bb;
            const rh_pair& p = sys.pairs[k];
            RH_REQUIRE(p.rc < 7.9, "cutoff must be below 8 sigma");
            d.pt.rc2[k] = (float)(p.rc * p.rc);
            d.pt.eps4[k] = (float)(4.0 * p.eps);
            d.pt.sigma2[k] = (float)(p.sigma * p.sigma);
            d.pt.shift[k] = (float)core::pair_shift(p);
            d.pt.six_eps4[k] = 6.0f * d.pt.eps4[k];
            double r = (p.rc + skin) * (1.0 + 1e-4);
            d.pt.list_r2[k] = (float)(r * r);
            d.pt.r2_fast[k] = fast_division_floor(d.pt.rc2[k], d.pt.sigma2[k], d.pt.six_eps4[k]);
        }
    double hs = 0.5 * skin * (1.0 - 1e-3);
    d.half_skin2 = (float)(hs * hs);
    d.face16 = face_margin16(d.pt, d.half_skin2);
    d.range_check = ((rcm + skin) * (1.0 + 1e-4) + skin >= 7.99) ? 1u : 0u;
    // The row-ordered copy for the 3D half list (variant 1), as the single-system engine plans it:
    // fine x bins a power-of-two fraction of the cell wide, each row padded with ghost bins that
    // cover the list radius plus one on a periodic x axis; x absolute in f32 needs x below 512.
    d.rows = 0u;
    d.cell_words = d.ncells;
    if (D == 3 && d.variant == 1 && !d.range_check && d.nc[0] >= 3 && d.nc[1] >= 3 && d.nc[2] >= 3 && b[0] <= 512.0) {
        float lr2max = 0.f;
        for (int k = 0; k < 16; k++) lr2max = std::max(lr2max, d.pt.list_r2[k]);
        const double reach = std::sqrt((double)lr2max) + 1e-3;
        const bool wrap_x = (d.periodic & 1u) != 0u;
        for (uint32_t sub = 8; sub >= 1 && !d.rows; sub /= 2) {
            const uint64_t nxf = (uint64_t)d.nc[0] * sub;
            const double wx = b[0] / (double)nxf;
            const uint64_t g = wrap_x ? (uint64_t)std::ceil(reach / wx) + 1u : 0u;
            const uint64_t words = (uint64_t)d.nc[1] * d.nc[2] * (nxf + 2 * g) + 1u;
            // a chord never holds a slot and its ghost, and no slot has two ghosts
            if (words > (uint64_t(1) << 25) || nxf <= 2 * g || !(b[0] > 2.0 * reach + 2.0 * wx)) continue;
            d.rows = 1u;
            d.rows_nxf = (uint32_t)nxf;
            d.rows_g = (uint32_t)g;
            d.rows_ne = (uint32_t)(nxf + 2 * g);
            d.cell_words = (uint32_t)(d.ncells + words);
        }
        if (d.rows) {
            d.rows_xscale = (double)d.rows_nxf / (double)sys.box_len[0];
            d.rows_inv_wx = (float)((double)d.rows_nxf / b[0]);
            d.rows_reach = (float)reach;
            d.bkind = 4u + (sys.nspecies > 1 ? 1u : 0u);
        }
    }
    const double pi = 3.141592653589793;
    double r = rcm + skin;
    double dens = (double)n / vol;
    double c = D == 3 ? std::ceil(4.0 / 3.0 * pi * r * r * r * std::fmax(2.0 * dens, 1.2)) : std::ceil(pi * r * r * std::fmax(2.0 * dens, 1.2));
    *cap = ((uint32_t)std::min(c + 8.0, 65532.0) + 3u) & ~3u;  // whole four-entry chunks
}

void Batch::set_tuning(int dim, uint32_t size_class, uint32_t variant, double cell_factor, uint32_t block) {
    RH_REQUIRE(dim == 2 || dim == 3, "dimension must be 2 or 3");
    RH_REQUIRE(size_class < 32, "size class is floor(log2 n), below 32");
    RH_REQUIRE(variant <= 2, "list variant must be 0 (full), 1 (half, built over the rows in 3D) or 2 (half, over the cell stencil)");
    RH_REQUIRE(cell_factor >= 1.0 && cell_factor <= 4.0, "cell factor must be in [1, 4]");
    RH_REQUIRE(block == 256 || block == 512 || block == 1024, "block must be 256, 512 or 1024");
    tuning_[dim - 2][size_class] = Tuning{variant, cell_factor, block};
}

uint64_t Batch::bytes_needed(int dim, uint64_t n, const SystemDesc& sys, double skin) const {
    BSys d;
    uint32_t cap;
    SystemDesc s2 = sys;
    s2.dim = dim;
    plan(s2, n, skin, d, &cap);
    uint64_t nslots = (n + d.block - 1) / d.block * d.block;
    return nslots * SLOT_BYTES + (uint64_t)cap * nslots * 4 + (uint64_t)d.cell_words * 12;
}

int Batch::add(const SystemDesc& sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, double skin, uint64_t steps) {
    RH_REQUIRE(steps < (uint64_t(1) << 32), "steps per run must fit 32 bits");
    const int D = sys.dim;
    for (uint64_t i = 0; i < n; i++) {
        RH_REQUIRE(species[i] < sys.nspecies, "species out of range");
        for (int a = 0; a < D; a++) RH_REQUIRE(pos[D * i + a] >= 0 && pos[D * i + a] < sys.box_len[a], "position outside the box");
    }
    if (free_ids_.empty()) return -1;
    BSys d;
    uint32_t cap;
    plan(sys, n, skin, d, &cap);
    // whole blocks, from a multiple of the block size (the step's work-groups are its blocks)
    uint64_t nslots = (n + d.block - 1) / d.block * d.block;
    uint64_t s0, l0, c0;
    uint64_t lunits = ((uint64_t)cap * nslots + LIST_UNIT - 1) / LIST_UNIT;
    uint64_t cunits = (d.cell_words + CELL_UNIT - 1) / CELL_UNIT;
    if (!slots_.alloc(nslots / BATCH_WG, &s0, d.block / BATCH_WG)) return -1;
    if (!lists_.alloc(lunits, &l0)) {
        slots_.free(s0, nslots / BATCH_WG);
        return -1;
    }
    if (!cells_.alloc(cunits, &c0)) {
        slots_.free(s0, nslots / BATCH_WG);
        lists_.free(l0, lunits);
        return -1;
    }
    int id = free_ids_.back();
    free_ids_.pop_back();
    d.active = 1;
    d.slot0 = (uint32_t)(s0 * BATCH_WG);
    d.nslots = (uint32_t)nslots;
    d.nbr0 = l0 * LIST_UNIT;
    d.cap = cap;
    d.cell0 = (uint32_t)(c0 * CELL_UNIT);
    BatchMember m;
    m.sys = sys;
    m.n = n;
    m.slot0 = d.slot0;
    m.nslots = d.nslots;
    m.cell0 = d.cell0;
    m.ncells = d.cell_words;
    m.cap = cap;
    m.nbr0 = d.nbr0;
    m.nbr_entries = lunits * LIST_UNIT;
    m.skin = skin;
    // Upload the state into parity 0 of the new range (nothing in flight touches it).  The
    // system joins only once its upload succeeded; a failure gives its ranges and id back.
    try {
        RH_HIP(hipStreamSynchronize(stream_));
        std::vector<double> px(n);
        std::vector<int32_t> vx(n);
        std::vector<uint32_t> ids(n);
        for (int a = 0; a < D; a++) {
            for (uint64_t i = 0; i < n; i++) px[i] = (double)pos[D * i + a];  // exact: |x| < 2^52
            RH_HIP(hipMemcpy(B_.pos[0][a] + d.slot0, px.data(), n * 8, hipMemcpyHostToDevice));
            for (uint64_t i = 0; i < n; i++) vx[i] = vel[D * i + a];
            RH_HIP(hipMemcpy(B_.vel[0][a] + d.slot0, vx.data(), n * 4, hipMemcpyHostToDevice));
        }
        RH_HIP(hipMemcpy(B_.species[0] + d.slot0, species, n * 4, hipMemcpyHostToDevice));
        for (uint64_t i = 0; i < n; i++) ids[i] = (uint32_t)i;
        RH_HIP(hipMemcpy(B_.ids[0] + d.slot0, ids.data(), n * 4, hipMemcpyHostToDevice));
        BCtl c;
        std::memset(&c, 0, sizeof c);
        c.stale = 1u;
        c.remaining = (uint32_t)steps;
        RH_HIP(hipMemcpy(d_ctl_ + id, &c, sizeof c, hipMemcpyHostToDevice));
        RH_HIP(hipMemset(B_.ledger + (size_t)id * LEDGER_WORDS, 0, sizeof(unsigned long long) * LEDGER_WORDS));
    } catch (...) {
        slots_.free(s0, nslots / BATCH_WG);
        lists_.free(l0, lunits);
        cells_.free(c0, cunits);
        free_ids_.push_back(id);
        throw;
    }
    members_[id] = m;
    h_sys_[id] = d;
    tables_dirty_ = true;
    return id;
}

void Batch::remove(int id) {
    auto it = members_.find(id);
    RH_REQUIRE(it != members_.end(), "no such system in the batch");
    RH_HIP(hipStreamSynchronize(stream_));
    const BatchMember& m = it->second;
    slots_.free(m.slot0 / BATCH_WG, m.nslots / BATCH_WG);
    lists_.free(m.nbr0 / LIST_UNIT, m.nbr_entries / LIST_UNIT);
    cells_.free(m.cell0 / CELL_UNIT, (m.ncells + CELL_UNIT - 1) / CELL_UNIT);
    h_sys_[id].active = 0;
    members_.erase(it);
    free_ids_.push_back(id);
    tables_dirty_ = true;
}

void Batch::set_steps(int id, uint64_t steps) {
    RH_REQUIRE(members_.count(id), "no such system in the batch");
    RH_REQUIRE(steps < (uint64_t(1) << 32), "steps per run must fit 32 bits");
    RH_HIP(hipStreamSynchronize(stream_));
    BCtl c;
    RH_HIP(hipMemcpy(&c, d_ctl_ + id, sizeof c, hipMemcpyDeviceToHost));
    c.remaining = (uint32_t)steps;
    RH_HIP(hipMemcpy(d_ctl_ + id, &c, sizeof c, hipMemcpyHostToDevice));
}

void Batch::upload_tables() {
    RH_HIP(hipStreamSynchronize(stream_));
    RH_HIP(hipMemcpy(d_sys_, h_sys_.data(), sizeof(BSys) * max_sys_, hipMemcpyHostToDevice));
    uint64_t hw = slots_.high_water();  // units of BATCH_WG
    // padded to whole blocks of the largest size: a step work-group reads its block's first unit
    constexpr uint32_t U = BATCH_MAX_BLOCK / BATCH_WG;
    std::vector<uint32_t> wg((std::max<uint64_t>(hw, 1) + U - 1) / U * U, ~0u);
    for (auto& kv : members_) {
        const BatchMember& m = kv.second;
        for (uint32_t w = m.slot0 / BATCH_WG; w < (m.slot0 + m.nslots) / BATCH_WG; w++) wg[w] = (uint32_t)kv.first;
    }
    RH_HIP(hipMemcpy(d_wg_, wg.data(), wg.size() * 4, hipMemcpyHostToDevice));
    uint32_t new_wgs = (uint32_t)std::max<uint64_t>(hw, 1);
    // Work lists: the rebuild lists' room per build kind (their systems' units), and per kind of
    // step launch (kind j = 0 (2D), 2 (3D), 4 (3D with a ledger), 6 (2D with a ledger), + (species
    // > 1); block 256 << b) the blocks of its systems.
    uint32_t rb_cap[RB_ALL + 1] = {};
    std::vector<uint32_t> lists[24];
    for (auto& kv : members_) {
        const BSys& d = h_sys_[kv.first];
        rb_cap[d.bkind] += d.nslots / BATCH_WG;
        rb_cap[RB_ALL] += d.nslots / BATCH_WG;
        const uint32_t j = (d.ledger ? (d.dim == 3 ? 4u : 6u) : (d.dim == 3 ? 2u : 0u)) + (d.nspecies > 1 ? 1u : 0u);
        const uint32_t bi = d.block == 1024u ? 2u : d.block == 512u ? 1u : 0u;
        for (uint32_t g = d.slot0 / d.block; g < (d.slot0 + d.nslots) / d.block; g++) lists[8 * bi + j].push_back(g);
    }
    std::vector<uint32_t> flat;
    bool changed = new_wgs != grid_wgs_;
    for (int k = 0; k < 24; k++) {
        changed |= step_cnt_[k] != (uint32_t)lists[k].size() || step_off_[k] != (uint32_t)flat.size();
        step_off_[k] = (uint32_t)flat.size();
        step_cnt_[k] = (uint32_t)lists[k].size();
        flat.insert(flat.end(), lists[k].begin(), lists[k].end());
    }
    if (!flat.empty()) RH_HIP(hipMemcpy(d_step_list_, flat.data(), flat.size() * 4, hipMemcpyHostToDevice));
    uint32_t off = 0;
    for (uint32_t k = 0; k < BUILD_KINDS; k++) {
        changed |= rb_cap_[k] != rb_cap[k] || rb_off_[k] != off;
        rb_off_[k] = off;
        rb_cap_[k] = rb_cap[k];
        off += rb_cap[k];
    }
    changed |= rb_cap_[RB_ALL] != rb_cap[RB_ALL];
    rb_off_[RB_ALL] = units_cap_;  // the list of all kinds takes the second half
    rb_cap_[RB_ALL] = rb_cap[RB_ALL];
    if (changed) invalidate_graph();
    grid_wgs_ = new_wgs;
    tables_dirty_ = false;
}

void Batch::invalidate_graph() {
    if (graph_) {
        RH_HIP(hipGraphExecDestroy(graph_));
        graph_ = nullptr;
    }
}

template <uint32_t BS>
void Batch::launch_steps(uint32_t b, hipStream_t st, int dim) {
    const uint32_t* cnt = step_cnt_ + 8 * b;
    const uint32_t* off = step_off_ + 8 * b;
    if (dim == 2) {
        if (cnt[0]) k_bb_step<2, false, false, BS><<<cnt[0], BS, 0, st>>>(B_, off[0]);
        if (cnt[1]) k_bb_step<2, true, false, BS><<<cnt[1], BS, 0, st>>>(B_, off[1]);
        if (cnt[6]) k_bb_step<2, false, true, BS><<<cnt[6], BS, 0, st>>>(B_, off[6]);
        if (cnt[7]) k_bb_step<2, true, true, BS><<<cnt[7], BS, 0, st>>>(B_, off[7]);
    } else {
        if (cnt[2]) k_bb_step<3, false, false, BS><<<cnt[2], BS, 0, st>>>(B_, off[2]);
        if (cnt[3]) k_bb_step<3, true, false, BS><<<cnt[3], BS, 0, st>>>(B_, off[3]);
        if (cnt[4]) k_bb_step<3, false, true, BS><<<cnt[4], BS, 0, st>>>(B_, off[4]);
        if (cnt[5]) k_bb_step<3, true, true, BS><<<cnt[5], BS, 0, st>>>(B_, off[5]);
    }
}

// Work-groups of a per-slot rebuild kernel: enough to fill the chip, at most one per unit.
static uint32_t rebuild_grid(uint32_t units) { return std::max<uint32_t>(1u, std::min<uint32_t>(units, 2048u)); }

void Batch::enqueue_iteration() {
    const uint32_t sb = blocks_for(max_sys_, 256);
    RbOff ro;
    for (uint32_t k = 0; k <= RB_ALL; k++) ro.off[k] = rb_off_[k];
    k_bb_decide<<<max_sys_, 64, 0, stream_>>>(d_sys_, d_ctl_, max_sys_, d_rb_list_, d_rb_count_, ro);
    const uint32_t ga = rebuild_grid(rb_cap_[RB_ALL]);
    k_bb_count<<<ga, BATCH_WG, 0, stream_>>>(B_, rb_off_[RB_ALL]);
    k_bb_scan<<<max_sys_, BATCH_WG, 0, stream_>>>(B_, max_sys_);
    k_bb_scatter<<<ga, BATCH_WG, 0, stream_>>>(B_, rb_off_[RB_ALL]);
    // From here on the dimensions go separate ways (they are different systems): on a side branch
    // the 3D systems' cell-stencil builds (those too small for rows), then the 2D systems' builds,
    // commit and steps; on the main stream the 3D row pipeline and builds, then (once the small
    // 3D systems' lists are built) the 3D commit and steps; the branches join before the step
    // commit.  In a captured graph the GPU runs the two side by side: the stencil builds are
    // latency-bound (long serial candidate loops in few waves) and the 3D step waits on its
    // gathers, so the work of one branch fills the other's stalls.
    const bool d2 = rb_cap_[0] + rb_cap_[1] != 0u, d3s = rb_cap_[2] + rb_cap_[3] != 0u, d3r = rb_cap_[4] + rb_cap_[5] != 0u;
    const bool d3 = d3s || d3r;
    // one side branch (a second hardware queue, no more: a serving process runs other streams
    // beside the batch): the small 3D systems' stencil builds first, then the 2D systems' work
    const bool fork = (d2 && d3) || (d3s && d3r);
    hipStream_t s2 = fork ? aux_ : stream_, s3 = d3s && d3r ? aux_ : stream_;
    if (fork) {
        RH_HIP(hipEventRecord(ev_fork_, stream_));
        RH_HIP(hipStreamWaitEvent(aux_, ev_fork_, 0));
    }
    if (rb_cap_[2]) k_bb_build<3, false><<<rebuild_grid(rb_cap_[2]), BATCH_WG, 0, s3>>>(B_, rb_off_[2]);
    if (rb_cap_[3]) k_bb_build<3, true><<<rebuild_grid(rb_cap_[3]), BATCH_WG, 0, s3>>>(B_, rb_off_[3]);
    if (s3 != stream_) RH_HIP(hipEventRecord(ev_join2_, aux_));  // the small 3D systems' lists are built
    if (d2) {
        if (rb_cap_[0]) k_bb_build<2, false><<<rebuild_grid(rb_cap_[0]), BATCH_WG, 0, s2>>>(B_, rb_off_[0]);
        if (rb_cap_[1]) k_bb_build<2, true><<<rebuild_grid(rb_cap_[1]), BATCH_WG, 0, s2>>>(B_, rb_off_[1]);
        k_bb_commit_rebuild<<<sb, 256, 0, s2>>>(d_sys_, d_ctl_, max_sys_, 2u);
        launch_steps<256>(0, s2, 2);
        launch_steps<512>(1, s2, 2);
        launch_steps<1024>(2, s2, 2);
    }
    // the row-ordered copies of the rebuilding rows systems, then their builds
    if (d3r) {
        for (uint32_t k = 4; k < 6; k++)
            if (rb_cap_[k]) k_bb_row_count<<<rebuild_grid(rb_cap_[k]), BATCH_WG, 0, stream_>>>(B_, rb_off_[k], k);
        k_bb_row_scan<<<max_sys_, ROW_SCAN_WG, 0, stream_>>>(B_, max_sys_);
        for (uint32_t k = 4; k < 6; k++)
            if (rb_cap_[k]) k_bb_row_scatter<<<rebuild_grid(rb_cap_[k]), BATCH_WG, 0, stream_>>>(B_, rb_off_[k], k);
        if (rb_cap_[4]) k_bb_build_rows<false><<<rebuild_grid(rb_cap_[4]), BATCH_WG, 0, stream_>>>(B_, rb_off_[4]);
        if (rb_cap_[5]) k_bb_build_rows<true><<<rebuild_grid(rb_cap_[5]), BATCH_WG, 0, stream_>>>(B_, rb_off_[5]);
    }
    if (s3 != stream_) RH_HIP(hipStreamWaitEvent(stream_, ev_join2_, 0));
    if (d3) {
        k_bb_commit_rebuild<<<sb, 256, 0, stream_>>>(d_sys_, d_ctl_, max_sys_, 3u);
        launch_steps<256>(0, stream_, 3);
        launch_steps<512>(1, stream_, 3);
        launch_steps<1024>(2, stream_, 3);
    }
    if (fork) {
        RH_HIP(hipEventRecord(ev_join_, aux_));
        RH_HIP(hipStreamWaitEvent(stream_, ev_join_, 0));
    }
    k_bb_commit_step<<<sb, 256, 0, stream_>>>(d_sys_, d_ctl_, max_sys_, d_rb_count_);
}

void Batch::run(uint64_t iterations) {
    if (tables_dirty_) upload_tables();
    if (members_.empty() || iterations == 0) return;
    uint64_t left = iterations;
    if (graph_chunk_ > 0 && iterations >= graph_chunk_) {
        if (!graph_ || graph_wgs_ != grid_wgs_ || graph_len_ != graph_chunk_) {
            invalidate_graph();
            hipGraph_t g;
            RH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (uint32_t k = 0; k < graph_chunk_; k++) enqueue_iteration();
            RH_HIP(hipStreamEndCapture(stream_, &g));
            RH_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            RH_HIP(hipGraphDestroy(g));
            graph_wgs_ = grid_wgs_;
            graph_len_ = graph_chunk_;
        }
        while (left >= graph_chunk_) {
            RH_HIP(hipGraphLaunch(graph_, stream_));
            left -= graph_chunk_;
        }
    }
    for (uint64_t k = 0; k < left; k++) enqueue_iteration();
    RH_HIP(hipGetLastError());
}

void Batch::sync() { RH_HIP(hipStreamSynchronize(stream_)); }

void Batch::handle_overflow() {
    for (auto& kv : members_) {
        int id = kv.first;
        BCtl& c = h_ctl_[id];
        if (!c.overflow) continue;
        BatchMember& m = kv.second;
        m.overflowed = true;
        uint32_t want = std::max<uint32_t>(c.maxcount + c.maxcount / 4 + 8, m.cap + m.cap / 2);
        want = (std::min<uint32_t>(want, 65532u) + 3u) & ~3u;
        uint64_t lunits = ((uint64_t)want * m.nslots + LIST_UNIT - 1) / LIST_UNIT;
        uint64_t l0;
        lists_.free(m.nbr0 / LIST_UNIT, m.nbr_entries / LIST_UNIT);
        if (!lists_.alloc(lunits, &l0)) {
            // No room to grow now: keep the old range; the system waits (its steps are gated).
            lists_.alloc(m.nbr_entries / LIST_UNIT, &l0);
            m.nbr0 = l0 * LIST_UNIT;
            h_sys_[id].nbr0 = m.nbr0;
            tables_dirty_ = true;
            continue;
        }
        m.cap = want;
        m.nbr0 = l0 * LIST_UNIT;
        m.nbr_entries = lunits * LIST_UNIT;
        h_sys_[id].cap = want;
        h_sys_[id].nbr0 = m.nbr0;
        c.overflow = 0u;
        c.stale = 1u;
        RH_HIP(hipMemcpy(d_ctl_ + id, &c, sizeof c, hipMemcpyHostToDevice));
        tables_dirty_ = true;
    }
}

void Batch::poll(uint64_t* remaining, uint32_t count) {
    RH_HIP(hipMemcpyAsync(h_ctl_, d_ctl_, sizeof(BCtl) * max_sys_, hipMemcpyDeviceToHost, stream_));
    RH_HIP(hipStreamSynchronize(stream_));
    handle_overflow();
    if (remaining)
        for (uint32_t i = 0; i < count && i < max_sys_; i++) remaining[i] = members_.count((int)i) ? h_ctl_[i].remaining : 0;
}

void Batch::read_ledger(int id, double energy[7], double impulse[7], uint64_t hits[7]) {
    RH_REQUIRE(members_.count(id), "no such system in the batch");
    RH_HIP(hipStreamSynchronize(stream_));
    unsigned long long L[LEDGER_WORDS];
    RH_HIP(hipMemcpy(L, B_.ledger + (size_t)id * LEDGER_WORDS, sizeof L, hipMemcpyDeviceToHost));
    for (int w = 0; w < 7; w++) {
        energy[w] = (double)(long long)L[3 * w] / LEDGER_ONE;
        impulse[w] = (double)(long long)L[3 * w + 1] / LEDGER_ONE;
        hits[w] = L[3 * w + 2];
    }
}

void Batch::read_state(int id, int64_t* pos, int32_t* vel, uint32_t* species) {
    auto it = members_.find(id);
    RH_REQUIRE(it != members_.end(), "no such system in the batch");
    const BatchMember& m = it->second;
    RH_HIP(hipStreamSynchronize(stream_));
    BCtl c;
    RH_HIP(hipMemcpy(&c, d_ctl_ + id, sizeof c, hipMemcpyDeviceToHost));
    const int D = m.sys.dim;
    const uint64_t n = m.n;
    std::vector<uint32_t> ids(n), sp(n);
    std::vector<double> px(n);
    std::vector<int32_t> vx(n);
    RH_HIP(hipMemcpy(ids.data(), B_.ids[c.pmeta] + m.slot0, n * 4, hipMemcpyDeviceToHost));
    for (int a = 0; a < D; a++) {
        RH_HIP(hipMemcpy(px.data(), B_.pos[c.ppos][a] + m.slot0, n * 8, hipMemcpyDeviceToHost));
        RH_HIP(hipMemcpy(vx.data(), B_.vel[c.pmeta][a] + m.slot0, n * 4, hipMemcpyDeviceToHost));
        for (uint64_t s = 0; s < n; s++) {
            if (pos) pos[D * (uint64_t)ids[s] + a] = (int64_t)px[s];
            if (vel) vel[D * (uint64_t)ids[s] + a] = vx[s];
        }
    }
    if (species) {
        RH_HIP(hipMemcpy(sp.data(), B_.species[c.pmeta] + m.slot0, n * 4, hipMemcpyDeviceToHost));
        for (uint64_t s = 0; s < n; s++) species[ids[s]] = sp[s];
    }
}

rh_list_stats Batch::stats(int id) {
    auto it = members_.find(id);
    RH_REQUIRE(it != members_.end(), "no such system in the batch");
    RH_HIP(hipStreamSynchronize(stream_));
    BCtl c;
    RH_HIP(hipMemcpy(&c, d_ctl_ + id, sizeof c, hipMemcpyDeviceToHost));
    rh_list_stats s;
    s.rebuilds = c.rebuilds;
    s.overflowed = c.overflow || it->second.overflowed ? 1u : 0u;
    s.cap = it->second.cap;
    s.max_count = c.maxcount;
    s.steps_done = ((uint64_t)c.done_hi << 32) | c.done_lo;
    s.launches_skipped = c.skipped;
    return s;
}

}  // namespace rh
