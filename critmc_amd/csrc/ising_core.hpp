// Bit-sliced word update shared by every Ising-type kernel (host-callable
// too, so the word logic can be exercised without a GPU).
#pragma once
#include "common.hpp"
#include "philox.hpp"

namespace cm {

constexpr int TAB_WORDS = 64;      // per replica: Metropolis table [0,32), heat bath [32,64)
constexpr int TAB_T = 15;          // offset of the thresholds inside one 32-word table
constexpr int NC_LIMIT = 14;       // (z+1) * 2 classes at z = 6

struct Sched {
    int32_t n;
    int32_t code[8];
    int32_t count[8];
};

struct Delta {
    int de, dm, acc;
};

__host__ __device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t& s,
                                         uint32_t& k)
{
    const uint32_t ab = a ^ b;
    s = ab ^ c;
    k = (a & b) | (c & ab);
}

// Bit-sliced population count of Z one-bit inputs per lane: count = c0 + 2 c1 + 4 c2.
template <int Z>
__host__ __device__ __forceinline__ void count_bits(const uint32_t (&in)[Z], uint32_t& c0, uint32_t& c1,
                                           uint32_t& c2)
{
    if constexpr (Z == 2) {
        c0 = in[0] ^ in[1];
        c1 = in[0] & in[1];
        c2 = 0;
    } else if constexpr (Z == 4) {
        uint32_t s1, k1;
        full_add(in[0], in[1], in[2], s1, k1);
        c0 = s1 ^ in[3];
        const uint32_t k2 = s1 & in[3];
        c1 = k1 ^ k2;
        c2 = k1 & k2;
    } else {
        uint32_t s1, k1, s2, k2;
        full_add(in[0], in[1], in[2], s1, k1);
        full_add(in[3], in[4], in[5], s2, k2);
        c0 = s1 ^ s2;
        const uint32_t k3 = s1 & s2;
        full_add(k1, k2, k3, c1, c2);
    }
}

__host__ __device__ __forceinline__ int popc(uint32_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __popc(v);
#else
    return __builtin_popcount(v);
#endif
}

// Lanes whose uniform is below their class threshold.  `tab` is one 32-word
// table: [0] = class count, [1 + k] = count value | spin selector << 8
// (0 any, 1 spin -1, 2 spin +1), [TAB_T + k] = threshold.  Lanes in no class
// accept unconditionally (p >= 1).
template <int NCMAX>
__host__ __device__ __forceinline__ uint32_t bitserial_below(uint32_t c0, uint32_t c1, uint32_t c2,
                                                    uint32_t own, uint32_t valid,
                                                    const uint32_t* __restrict__ tab, uint32_t k0,
                                                    uint32_t k1, uint32_t sid, uint64_t sweep,
                                                    uint32_t colour, uint32_t group)
{
    const uint32_t ncls = tab[0];
    uint32_t mask[NCMAX];
    uint32_t thr[NCMAX];
    uint32_t pending = 0;
#pragma unroll
    for (int k = 0; k < NCMAX; ++k) {
        uint32_t mk = 0;
        if ((uint32_t)k < ncls) {
            const uint32_t d = tab[1 + k];
            const uint32_t v = d & 0xffu, sel = d >> 8;
            mk = (c0 ^ ((v & 1u) ? 0u : ~0u)) & (c1 ^ ((v & 2u) ? 0u : ~0u)) &
                 (c2 ^ ((v & 4u) ? 0u : ~0u));
            if (sel == 1u) mk &= ~own;
            else if (sel == 2u) mk &= own;
            mk &= valid;
        }
        mask[k] = mk;
        thr[k] = tab[TAB_T + k];
        pending |= mk;
    }
    uint32_t below = valid & ~pending;
    const uint32_t zc = ctr_z(sweep, colour, TAG_ISING);
    for (uint32_t q = 0; q < 8u && pending != 0u; ++q) {
        const u32x4 r = philox4x32_10(u32x4{8u * group + q, (uint32_t)sweep, zc, sid}, k0, k1);
        const uint32_t planes[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t shift = 31u - (4u * q + (uint32_t)t);
            uint32_t tb = 0;
#pragma unroll
            for (int k = 0; k < NCMAX; ++k) tb |= mask[k] & (0u - ((thr[k] >> shift) & 1u));
            const uint32_t diff = pending & (planes[t] ^ tb);
            below |= diff & tb;
            pending &= ~diff;
        }
    }
    return below;
}

// One word of 32 same-colour sites; returns the new spins and accumulates deltas.
template <int Z, int NCMAX>
__host__ __device__ __forceinline__ uint32_t update_word(int code, uint32_t own, const uint32_t (&nb)[Z],
                                                uint32_t valid, const uint32_t* __restrict__ tabs,
                                                uint32_t k0, uint32_t k1, uint32_t sid,
                                                uint64_t sweep, uint32_t colour, uint32_t group,
                                                Delta& d)
{
    uint32_t in[Z];
    uint32_t c0, c1, c2;
    if (code == CODE_METROPOLIS) {
#pragma unroll
        for (int k = 0; k < Z; ++k) in[k] = ~(own ^ nb[k]);  // aligned neighbours
        count_bits<Z>(in, c0, c1, c2);
        const uint32_t F = bitserial_below<NCMAX>(c0, c1, c2, own, valid, tabs, k0, k1, sid,
                                                  sweep, colour, group) & valid;
        const int nf = popc(F);
        const int nal = popc(F & c0) + 2 * popc(F & c1) + 4 * popc(F & c2);
        d.de += 4 * nal - 2 * Z * nf;          // sum of 2a, a = 2 n_aligned - z
        d.dm -= 2 * (2 * popc(F & own) - nf);  // sum of -2 s_old
        d.acc += nf;
        return own ^ F;
    } else {
        count_bits<Z>(nb, c0, c1, c2);  // neighbours up
        const uint32_t up = bitserial_below<NCMAX>(c0, c1, c2, own, valid, tabs + 32, k0, k1,
                                                   sid, sweep, colour, group);
        const uint32_t X = (up ^ own) & valid;
        const uint32_t Xu = X & own, Xd = X & ~own;
        const int fu = 2 * (popc(Xu & c0) + 2 * popc(Xu & c1) + 4 * popc(Xu & c2)) - Z * popc(Xu);
        const int fd = 2 * (popc(Xd & c0) + 2 * popc(Xd & c1) + 4 * popc(Xd & c2)) - Z * popc(Xd);
        d.de += 2 * (fu - fd);                 // sum of 2 s_old f
        d.dm += 2 * (popc(Xd) - popc(Xu));     // sum of 2 s_new
        d.acc += popc(X);
        return own ^ X;
    }
}

}  // namespace cm
