// gpu-prims for gfx950: exclusive scan, stable LSD radix sort, reduction, fills, Threefry.
// The MI355X counterparts of the reference's crates/gpu-prims (ScanPlan prims.rs:157-196,
// BinPlan prims.rs:198-248, ReducePlan prims.rs:116-155, Gated prims.rs:250-283,
// threefry.rs:3-46).  Every launch takes an optional device gate word (see rh_common.h), the
// equivalent of the reference's indirect-dispatch gate: work decided on the device never
// round-trips to the host.
#pragma once

#include "rh_common.h"

namespace rh {

// Temporary bytes an exclusive scan of n u32 needs.
size_t scan_temp_bytes(uint64_t n);
// out[i] = sum(in[0..i)), u32 wrapping.  in == out is allowed.
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* temp, const uint32_t* gate, hipStream_t s);

// A stable LSD radix sort of (key, value) u32 pairs over the low `key_bits` bits, 8 bits a pass.
struct RadixSort {
    uint64_t cap = 0;  // capacity in items
    uint32_t nblocks_cap = 0;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;  // 256 * nblocks
    void* scan_temp = nullptr;

    void reserve(uint64_t n);
    void release();
    uint64_t bytes() const;  // device memory held
    ~RadixSort() { release(); }
    // Sorts keys[0]/vals[0] (n items); returns the index (0 or 1) of the buffers holding the
    // result.  The number of passes is ceil(key_bits / 8), so the result index is known on the
    // host without a readback.
    int sort(uint64_t n, uint32_t key_bits, const uint32_t* gate, hipStream_t s);
    static int result_index(uint32_t key_bits) { return (int)(((key_bits + 7) / 8) & 1u); }
};

// The same sort with no allocation: the caller's arrays are one buffer of the ping-pong pair and
// a fixed device array of the library holds the rest (at most SORT_FIXED_MAX items).  Enqueued
// on s without synchronising.  Calls that overlap on one GPU share that array and race.
constexpr uint64_t SORT_FIXED_MAX = 1ull << 20;
void radix_sort_u32_fixed(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits, hipStream_t s);

// Sum of n doubles into *out (a fixed-shape tree, so the result is run-to-run identical).
size_t reduce_temp_bytes(uint64_t n);
void reduce_sum_f64(const double* in, double* out, uint64_t n, void* temp, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, const uint32_t* gate, hipStream_t s);

// Threefry-2x32-20 on the device: out[2i], out[2i+1] = threefry((i + ctr0, ctr1), key).
void threefry_fill(uint32_t* out, uint64_t n, uint32_t ctr0, uint32_t ctr1, uint32_t k0, uint32_t k1, hipStream_t s);

// Two-word (64-bit) atomic accumulation check: adds vals[i] (i64) into *acc with native 64-bit
// atomics.
void atomic_add_i64(const int64_t* vals, uint64_t n, int64_t* acc, hipStream_t s);

__host__ __device__ __forceinline__ void threefry2x32(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, uint32_t out[2]) {
    const uint32_t rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
    uint32_t ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
    uint32_t x0 = c0 + ks[0], x1 = c1 + ks[1];
#pragma unroll
    for (uint32_t r = 0; r < 20; r++) {
        x0 += x1;
        x1 = (x1 << rot[r % 8]) | (x1 >> (32 - rot[r % 8]));
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

}  // namespace rh
