// Shared definitions of the critmc_amd HIP library (gfx950).
//
// Random stream contract (every kernel follows it; the CPU replay restates it):
//   key = (k0, k1) per chain
//   ctr = (block, sweep_lo, sweep_hi16 | colour << 16 | tag << 24, stream_id)
// `sweep` is the chain's own count of local sweeps, so a run split into any
// number of calls draws the same numbers (critmc tests/test_sim_chunks.py).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

namespace cm {

enum Tag : uint32_t {
    TAG_INIT = 1,
    TAG_ISING = 2,
    TAG_POTTS = 3,
    TAG_EA = 4,
    TAG_PT = 5,
    TAG_ON = 6,
    TAG_ONHB = 8,  // O(n) heat bath (tag 7 belongs to the CPU typewriter baseline)
};

// Update codes of critmc/kernels/common.py:16-21.
enum Code : int64_t {
    CODE_METROPOLIS = 0,
    CODE_METROPOLIS_RANDOM = 1,
    CODE_HEATBATH = 2,
    CODE_WOLFF = 3,
    CODE_SWENDSEN_WANG = 4,
    CODE_OVERRELAX = 5,
};

// Return codes of every C entry point.
enum Status : int {
    CM_OK = 0,
    CM_EINVAL = -1,     // bad argument (shape, size, code)
    CM_EUNSUPPORTED = -2,
    CM_EHIP = -3,       // a HIP runtime call failed
};

void set_error(const char* fmt, ...);
const char* last_error();

// Stream-ordered host-to-device copy of a small host buffer.  The bytes go
// through a pinned staging buffer owned by the calling thread, so `src` may be
// released as soon as this returns; the staging buffer is reused only after
// the previous copy from it has completed.
hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t stream);

#define CM_HIP_CHECK(expr)                                                              \
    do {                                                                                \
        hipError_t err__ = (expr);                                                      \
        if (err__ != hipSuccess) {                                                      \
            ::cm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(err__),     \
                            __FILE__, __LINE__);                                        \
            return ::cm::CM_EHIP;                                                       \
        }                                                                               \
    } while (0)

#define CM_REQUIRE(cond, ...)                                                           \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::cm::set_error(__VA_ARGS__);                                               \
            return ::cm::CM_EINVAL;                                                     \
        }                                                                               \
    } while (0)

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its own
// L2.  xcd_block maps the dispatch index w of a grid of `total` workgroups to a
// logical block so that each XCD receives one contiguous range of logical
// blocks; lattice rows processed by neighbouring blocks then share an L2.  The
// map is a bijection for any dispatch order, so it never changes results.
__host__ __device__ __forceinline__ uint32_t xcd_block(uint32_t w, uint32_t total)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t q = total / kXcd, r = total % kXcd;
    const uint32_t x = w % kXcd, j = w / kXcd;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Integer sum over the wave with DPP adds: inclusive sums within each row of
// 16 lanes (row_shr 1, 2, 4, 8, lanes shifted in from outside the row read 0),
// then row 0's and row 2's totals carried into rows 1 and 3 (row_bcast:15)
// and the first half's total into the second half (row_bcast:31); lane 63
// then holds the total, returned to every lane.  Six adds instead of six
// shuffles through LDS.
__device__ __forceinline__ int wave_sum_dpp(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

// Sum over the wave of a per-lane value in [0, 2^B): one ballot and popcount
// per bit instead of a shuffle tree (the result is wave-uniform).
template <int B>
__device__ __forceinline__ int wave_sum_bits(uint32_t v)
{
    int s = 0;
#pragma unroll
    for (int b = 0; b < B; ++b) s += __popcll(__ballot((v >> b) & 1u)) << b;
    return s;
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace cm
