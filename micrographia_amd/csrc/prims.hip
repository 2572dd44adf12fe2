// gpu-prims for gfx950 (see prims.h).  Wave64 scans use cross-lane shuffles, the radix sort ranks
// digits with ballots (a wave-level match over the digit's 8 bits) and stages per-block digit
// histograms in LDS.
#include "prims.h"

#include <mutex>
#include <string>

namespace rh {

// ---------------------------------------------------------------------------------------------
// Exclusive scan (u32): 256 threads x 4 items per block, block totals scanned recursively.
// ---------------------------------------------------------------------------------------------

constexpr int SCAN_T = 256;
constexpr int SCAN_IPT = 4;
constexpr int SCAN_ITEMS = SCAN_T * SCAN_IPT;

// Exclusive block scan of one value per thread; *total gets the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_wave, uint32_t* total) {
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t y = __shfl_up(x, d, WAVE);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == WAVE - 1) lds_wave[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN_T / WAVE; k++) {
        uint32_t s = lds_wave[k];
        pre += (k < (int)w) ? s : 0u;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return pre + x - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_partial(const uint32_t* in, uint32_t* out, uint32_t* sums, uint64_t n, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    __shared__ uint32_t lds_wave[SCAN_T / WAVE];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_ITEMS + (uint64_t)threadIdx.x * SCAN_IPT;
    uint32_t a[SCAN_IPT];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        a[k] = (base + k < n) ? in[base + k] : 0u;
        s += a[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, lds_wave, &total);
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += a[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_add(uint32_t* out, const uint32_t* sums, uint64_t n, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    uint32_t add = sums[blockIdx.x];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_ITEMS + (uint64_t)threadIdx.x * SCAN_IPT;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++)
        if (base + k < n) out[base + k] += add;
}

static uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

size_t scan_temp_bytes(uint64_t n) {
    size_t bytes = 0;
    while (n > 1) {
        uint64_t nb = (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
        bytes += round_up(nb, 64) * sizeof(uint32_t);
        n = nb;
    }
    return bytes + 256;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, const uint32_t* gate, hipStream_t s) {
    if (n == 0) return;
    uint32_t nb = blocks_for(n, SCAN_ITEMS);
    uint32_t* sums = (uint32_t*)temp;
    k_scan_partial<<<nb, SCAN_T, 0, s>>>(in, out, sums, n, gate);
    RH_HIP(hipGetLastError());
    if (nb > 1) {
        exclusive_scan_u32(sums, sums, nb, sums + round_up(nb, 64), gate, s);
        k_scan_add<<<nb, SCAN_T, 0, s>>>(out, sums, n, gate);
        RH_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Block = 256 threads x 16 rounds = 4096 items.
// Pass: (1) per-block digit histogram in LDS -> global hist[digit * nblocks + block]
//       (2) exclusive scan of hist (digit-major, so it yields each (digit, block) offset)
//       (3) scatter: items ranked in order within the block (round by round; within a round a
//           ballot match gives each lane its rank among equal digits of its wave, and per-wave
//           digit counts in LDS give the offset of earlier waves).
// ---------------------------------------------------------------------------------------------

constexpr int RS_T = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_ITEMS = RS_T * RS_ROUNDS;
constexpr int RS_WAVES = RS_T / WAVE;

__global__ __launch_bounds__(RS_T) void k_radix_hist(const uint32_t* keys, uint64_t n, uint32_t shift, uint32_t* hist, uint32_t nblocks, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_ITEMS;
    for (int r = 0; r < RS_ROUNDS; r++) {
        uint64_t idx = base + (uint64_t)r * RS_T + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_T) void k_radix_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n, uint32_t shift,
                                                        const uint32_t* offsets, uint32_t nblocks, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    __shared__ uint32_t base[256];
    __shared__ uint32_t wc[RS_WAVES][256];
    const uint32_t t = threadIdx.x, w = t / WAVE;
    base[t] = offsets[(uint64_t)t * nblocks + blockIdx.x];
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) wc[k][t] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * RS_ITEMS;
    const uint64_t lt = lanemask_lt();
    for (int r = 0; r < RS_ROUNDS; r++) {
        uint64_t idx = b0 + (uint64_t)r * RS_T + t;
        bool valid = idx < n;
        uint32_t k = valid ? kin[idx] : 0u;
        uint32_t v = valid ? vin[idx] : 0u;
        uint32_t d = (k >> shift) & 255u;
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            uint64_t bb = ballot64((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) wc[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d] + rank;
            for (uint32_t q = 0; q < w; q++) off += wc[q][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int q = 0; q < RS_WAVES; q++) {
            add += wc[q][t];
            wc[q][t] = 0;
        }
        base[t] += add;
        __syncthreads();
    }
}

void RadixSort::reserve(uint64_t n) {
    if (n <= cap && keys[0]) return;
    release();
    cap = n ? n : 1;
    nblocks_cap = blocks_for(cap, RS_ITEMS);
    for (int b = 0; b < 2; b++) {
        RH_HIP(hipMalloc(&keys[b], cap * sizeof(uint32_t)));
        RH_HIP(hipMalloc(&vals[b], cap * sizeof(uint32_t)));
    }
    uint64_t hn = 256ull * nblocks_cap;
    RH_HIP(hipMalloc(&hist, hn * sizeof(uint32_t)));
    RH_HIP(hipMalloc(&scan_temp, scan_temp_bytes(hn)));
}

uint64_t RadixSort::bytes() const {
    if (!keys[0]) return 0;
    const uint64_t hn = 256ull * nblocks_cap;
    return 4 * cap * sizeof(uint32_t) + hn * sizeof(uint32_t) + scan_temp_bytes(hn);
}

void RadixSort::release() {
    for (int b = 0; b < 2; b++) {
        if (keys[b]) (void)hipFree(keys[b]);
        if (vals[b]) (void)hipFree(vals[b]);
        keys[b] = vals[b] = nullptr;
    }
    if (hist) (void)hipFree(hist);
    if (scan_temp) (void)hipFree(scan_temp);
    hist = nullptr;
    scan_temp = nullptr;
    cap = 0;
}

// The passes over a (keys, vals) ping-pong pair; hist holds 256 x blocks_for(n, RS_ITEMS) words
// and scan_temp scan_temp_bytes(256 x that).  Returns the index of the pair holding the result.
static int radix_sort_passes(uint32_t* const keys[2], uint32_t* const vals[2], uint32_t* hist, void* scan_temp, uint64_t n, uint32_t key_bits,
                             const uint32_t* gate, hipStream_t s) {
    uint32_t passes = (key_bits + 7) / 8;
    uint32_t nb = blocks_for(n, RS_ITEMS);
    int cur = 0;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t shift = 8 * p;
        k_radix_hist<<<nb, RS_T, 0, s>>>(keys[cur], n, shift, hist, nb, gate);
        RH_HIP(hipGetLastError());
        exclusive_scan_u32(hist, hist, 256ull * nb, scan_temp, gate, s);
        k_radix_scatter<<<nb, RS_T, 0, s>>>(keys[cur], vals[cur], keys[cur ^ 1], vals[cur ^ 1], n, shift, hist, nb, gate);
        RH_HIP(hipGetLastError());
        cur ^= 1;
    }
    return cur;
}

int RadixSort::sort(uint64_t n, uint32_t key_bits, const uint32_t* gate, hipStream_t s) {
    RH_REQUIRE(n <= cap, "radix sort: more items than reserved");
    return radix_sort_passes(keys, vals, hist, scan_temp, n, key_bits, gate, s);
}

// ---------------------------------------------------------------------------------------------
// The same sort without allocation.  Its scratch is one file-scope device array: the second
// key/value buffer of the ping-pong pair (the caller's arrays are the first), the digit
// histograms and the scan's block sums, sized for SORT_FIXED_MAX items.  Every GPU holds its own
// copy, at its own address, which hipGetSymbolAddress gives once per device; host code uses that
// address, never the array's name.  The histograms and block sums are written in full before
// they are read, so nothing in the array needs clearing between calls.
// ---------------------------------------------------------------------------------------------

constexpr uint64_t SORT_FIXED_BLOCKS = (SORT_FIXED_MAX + RS_ITEMS - 1) / RS_ITEMS;
constexpr size_t SORT_FIXED_HIST_WORDS = 256 * SORT_FIXED_BLOCKS;
constexpr size_t SORT_FIXED_SCAN_BYTES = 4096;
constexpr size_t SORT_FIXED_BYTES = 2 * SORT_FIXED_MAX * sizeof(uint32_t) + SORT_FIXED_HIST_WORDS * sizeof(uint32_t) + SORT_FIXED_SCAN_BYTES;

static __device__ __align__(256) uint8_t scratch[SORT_FIXED_BYTES];

constexpr int SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SCRATCH_DEVICES];

// The device address of `scratch` on the current device, looked up on the first call there.
static uint8_t* scratch_on_current_device() {
    int dev = -1;
    RH_HIP(hipGetDevice(&dev));
    RH_REQUIRE(dev >= 0 && dev < SCRATCH_DEVICES, "fixed-scratch radix sort: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        RH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RH_REQUIRE(p != nullptr, "fixed-scratch radix sort: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<uint8_t*>(g_scratch_ptr[dev]);
}

__global__ void k_copy_u32x2(const uint32_t* a, uint32_t* a_out, const uint32_t* b, uint32_t* b_out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        a_out[i] = a[i];
        b_out[i] = b[i];
    }
}

void radix_sort_u32_fixed(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits, hipStream_t s) {
    RH_REQUIRE(key_bits >= 1 && key_bits <= 32, "key_bits must be 1..32");
    RH_REQUIRE(n <= SORT_FIXED_MAX, "fixed-scratch radix sort: n = " + std::to_string(n) + " exceeds its fixed scratch of " +
                                        std::to_string(SORT_FIXED_MAX) + " items (use rh_prims_radix_sort_u32)");
    if (n == 0) return;
    RH_REQUIRE(keys != nullptr && vals != nullptr, "null key or value array");
    const uint32_t nb = blocks_for(n, RS_ITEMS);
    RH_REQUIRE(scan_temp_bytes(256ull * nb) <= SORT_FIXED_SCAN_BYTES, "fixed-scratch radix sort: scan scratch too small");
    uint8_t* base = scratch_on_current_device();
    uint32_t* k1 = reinterpret_cast<uint32_t*>(base);
    uint32_t* v1 = k1 + SORT_FIXED_MAX;
    uint32_t* hist = v1 + SORT_FIXED_MAX;
    void* scan_temp = hist + SORT_FIXED_HIST_WORDS;
    uint32_t* const ks[2] = {keys, k1};
    uint32_t* const vs[2] = {vals, v1};
    if (radix_sort_passes(ks, vs, hist, scan_temp, n, key_bits, nullptr, s) == 1) {
        uint32_t cb = blocks_for(n, 256);
        if (cb > 1024) cb = 1024;
        k_copy_u32x2<<<cb, 256, 0, s>>>(k1, keys, v1, vals, n);
        RH_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------------------------
// Reduction (f64 sum), fixed tree: 256 threads x 8 items per block.
// ---------------------------------------------------------------------------------------------

constexpr int RD_T = 256, RD_IPT = 8, RD_ITEMS = RD_T * RD_IPT;

__global__ __launch_bounds__(RD_T) void k_reduce_f64(const double* in, double* out, uint64_t n) {
    __shared__ double lds[RD_T];
    uint64_t base = (uint64_t)blockIdx.x * RD_ITEMS;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < RD_IPT; k++) {
        uint64_t idx = base + (uint64_t)k * RD_T + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int w = RD_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = lds[0];
}

size_t reduce_temp_bytes(uint64_t n) {
    size_t bytes = 0;
    while (n > 1) {
        n = (n + RD_ITEMS - 1) / RD_ITEMS;
        bytes += round_up(n, 32) * sizeof(double);
    }
    return bytes + 256;
}

void reduce_sum_f64(const double* in, double* out, uint64_t n, void* temp, hipStream_t s) {
    const double* cur = in;
    double* t = (double*)temp;
    if (n <= 1) {
        if (n == 1) RH_HIP(hipMemcpyAsync(out, in, sizeof(double), hipMemcpyDeviceToDevice, s));
        return;
    }
    while (n > 1) {
        uint32_t nb = blocks_for(n, RD_ITEMS);
        double* dst = nb == 1 ? out : t;
        k_reduce_f64<<<nb, RD_T, 0, s>>>(cur, dst, n);
        RH_HIP(hipGetLastError());
        cur = dst;
        t += round_up(nb, 32);
        n = nb;
    }
}

// ---------------------------------------------------------------------------------------------
// Fills, Threefry, 64-bit atomics
// ---------------------------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n, const uint32_t* gate) {
    if (gate_closed(gate)) return;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, const uint32_t* gate, hipStream_t s) {
    if (n == 0) return;
    uint32_t nb = blocks_for(n, 256);
    if (nb > 4096) nb = 4096;
    k_fill_u32<<<nb, 256, 0, s>>>(p, v, n, gate);
    RH_HIP(hipGetLastError());
}

__global__ void k_threefry(uint32_t* out, uint64_t n, uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t r[2];
    threefry2x32((uint32_t)i + c0, c1, k0, k1, r);
    out[2 * i] = r[0];
    out[2 * i + 1] = r[1];
}

void threefry_fill(uint32_t* out, uint64_t n, uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, hipStream_t s) {
    if (n == 0) return;
    k_threefry<<<blocks_for(n, 256), 256, 0, s>>>(out, n, c0, c1, k0, k1);
    RH_HIP(hipGetLastError());
}

__global__ void k_atomic_add_i64(const int64_t* vals, uint64_t n, int64_t* acc) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd((unsigned long long*)acc, (unsigned long long)vals[i]);
}

void atomic_add_i64(const int64_t* vals, uint64_t n, int64_t* acc, hipStream_t s) {
    if (n == 0) return;
    k_atomic_add_i64<<<blocks_for(n, 256), 256, 0, s>>>(vals, n, acc);
    RH_HIP(hipGetLastError());
}

}  // namespace rh
