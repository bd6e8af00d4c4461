// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) for device and host.
// Counter-based: any (key, counter) pair is evaluated independently, so the
// kernels draw exactly the numbers they need, in any order, on any thread.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cm {

struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Third counter word: high bits of the sweep, colour and purpose tag.
__host__ __device__ __forceinline__ uint32_t ctr_z(uint64_t sweep, uint32_t colour, uint32_t tag)
{
    return (uint32_t)((sweep >> 32) & 0xFFFFu) | (colour << 16) | (tag << 24);
}

__host__ __device__ __forceinline__ u32x4 draw(uint32_t k0, uint32_t k1, uint32_t stream,
                                               uint64_t sweep, uint32_t colour, uint32_t tag,
                                               uint32_t block)
{
    return philox4x32_10(u32x4{block, (uint32_t)sweep, ctr_z(sweep, colour, tag), stream}, k0, k1);
}

}  // namespace cm
