// Shared host/device helpers of the Rahman HIP engine (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace rh {

// An error raised inside the library; the C ABI turns it into a status code and rh_last_error().
struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define RH_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t rh_err_ = (call);                                                                  \
        if (rh_err_ != hipSuccess)                                                                    \
            throw ::rh::Error(std::string("HIP error in ") + #call + ": " + hipGetErrorString(rh_err_)); \
    } while (0)

#define RH_REQUIRE(cond, msg)                    \
    do {                                         \
        if (!(cond)) throw ::rh::Error(msg);     \
    } while (0)

constexpr int WAVE = 64;

inline uint32_t blocks_for(uint64_t n, uint32_t per_block) {
    uint64_t b = (n + per_block - 1) / per_block;
    return (uint32_t)(b ? b : 1);
}

// ---- device helpers -------------------------------------------------------------------------

// Float -> i32 with the hardware conversion: truncation, saturating at the i32 range, NaN -> 0.
// These are exactly Rust's `f as i32` semantics, which the reference CPU engine uses.
__device__ __forceinline__ int32_t cvt_sat_i32(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// cvt_sat_i32(-x), the negation taken by the instruction's source modifier.
__device__ __forceinline__ int32_t cvt_sat_i32_neg(float x) {
    int32_t r;
    asm("v_cvt_i32_f32_e64 %0, -%1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Ballot over the wave (64 lanes).
__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// Gate shared by gated launches: a null gate always runs; otherwise the kernel runs only if the
// word is non-zero.  All work-groups read the word before doing anything, and nothing in the
// gated kernel writes it, so a launch runs entirely or not at all.
__device__ __forceinline__ bool gate_closed(const uint32_t* gate) { return gate != nullptr && *(volatile const uint32_t*)gate == 0u; }

// The dispatcher hands work-groups to the chip's 8 XCDs round-robin (hardware id mod 8), and each
// XCD has its own L2.  Renumbering so that XCD x runs one contiguous range of logical blocks puts
// blocks that are neighbours in cell order, and so share most of their neighbour atoms, behind the
// same L2.  A bijection of [0, nblk) for any nblk.
__device__ __forceinline__ uint32_t xcd_block(uint32_t hw, uint32_t nblk) {
    constexpr uint32_t XCDS = 8;
    const uint32_t q = nblk / XCDS, r = nblk % XCDS, x = hw % XCDS, k = hw / XCDS;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace rh
