// Checkerboard geometry of periodic hypercubic lattices (d <= 3, even extents).
//
// Site numbering is x-fastest, i = x0 + L0 (x1 + L1 x2) (critmc/lattice.py:118-128).
// The colour-c sites are enumerated row by row: k = r * M + m with row
// r = x1 + L1 x2, M = L0/2 and x0 = 2m + ((x1 + x2 + c) & 1).  Neighbours are
// returned in the reference's column order +x, -x, +y, -y, +z, -z
// (critmc/lattice.py:129-134).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cm {

// Division of x < 2^31 by a runtime constant d >= 1 with one high multiply:
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1, q = (mulhi(x, m) + x) >> s
// (Granlund & Montgomery, PLDI 1994, the round-up variant).  x < 2^31 keeps the
// sum below 2^32.
struct FastDiv {
    uint32_t d, m, s;
};

__host__ __device__ inline FastDiv make_fastdiv(uint32_t d)
{
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    const uint64_t m = (((1ull << s) - d) << 32) / d + 1;
    return FastDiv{d, (uint32_t)m, s};
}

__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t x, const FastDiv& f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t hi = __umulhi(x, f.m);
#else
    const uint32_t hi = (uint32_t)(((uint64_t)x * f.m) >> 32);
#endif
    return (hi + x) >> f.s;
}

struct CbGeo {
    int d, L0, L1, L2, M;
    int64_t N, half;  // sites, sites per colour
    FastDiv divM, divL1;
};

__host__ __device__ inline void cb_finish(CbGeo& g)
{
    g.divM = make_fastdiv((uint32_t)g.M);
    g.divL1 = make_fastdiv((uint32_t)g.L1);
}

template <int Z>
__host__ __device__ __forceinline__ int64_t cb_site(const CbGeo& g, int64_t k, int c,
                                                    int64_t (&nb)[Z])
{
    const uint32_t r = fdiv((uint32_t)k, g.divM);
    const int m = (int)((uint32_t)k - r * (uint32_t)g.M);
    int x1 = 0, x2 = 0;
    if (Z >= 4) {
        const uint32_t q = fdiv(r, g.divL1);
        x1 = (int)(r - q * (uint32_t)g.L1);
        x2 = (int)q;
    }
    const int x0 = 2 * m + ((x1 + x2 + c) & 1);
    const int64_t base = (int64_t)r * g.L0;
    const int64_t i = base + x0;
    nb[0] = base + (x0 + 1 == g.L0 ? 0 : x0 + 1);
    nb[1] = base + (x0 == 0 ? g.L0 - 1 : x0 - 1);
    if (Z >= 4) {
        const int64_t row0 = (int64_t)x2 * g.L1;
        nb[2] = (row0 + (x1 + 1 == g.L1 ? 0 : x1 + 1)) * g.L0 + x0;
        nb[3] = (row0 + (x1 == 0 ? g.L1 - 1 : x1 - 1)) * g.L0 + x0;
    }
    if (Z == 6) {
        nb[4] = ((int64_t)(x2 + 1 == g.L2 ? 0 : x2 + 1) * g.L1 + x1) * g.L0 + x0;
        nb[5] = ((int64_t)(x2 == 0 ? g.L2 - 1 : x2 - 1) * g.L1 + x1) * g.L0 + x0;
    }
    return i;
}

// cb_site with 32-bit site indices, for lattices of fewer than 2^31 sites.
template <int Z>
__host__ __device__ __forceinline__ uint32_t cb_site32(const CbGeo& g, uint32_t k, int c,
                                                       uint32_t (&nb)[Z])
{
    const uint32_t r = fdiv(k, g.divM);
    const int m = (int)(k - r * (uint32_t)g.M);
    int x1 = 0, x2 = 0;
    if (Z >= 4) {
        const uint32_t q = fdiv(r, g.divL1);
        x1 = (int)(r - q * (uint32_t)g.L1);
        x2 = (int)q;
    }
    const int x0 = 2 * m + ((x1 + x2 + c) & 1);
    const uint32_t L0 = (uint32_t)g.L0, L1 = (uint32_t)g.L1;
    const uint32_t base = r * L0;
    nb[0] = base + (uint32_t)(x0 + 1 == g.L0 ? 0 : x0 + 1);
    nb[1] = base + (uint32_t)(x0 == 0 ? g.L0 - 1 : x0 - 1);
    if (Z >= 4) {
        const uint32_t row0 = (uint32_t)x2 * L1;
        nb[2] = (row0 + (uint32_t)(x1 + 1 == g.L1 ? 0 : x1 + 1)) * L0 + (uint32_t)x0;
        nb[3] = (row0 + (uint32_t)(x1 == 0 ? g.L1 - 1 : x1 - 1)) * L0 + (uint32_t)x0;
    }
    if (Z == 6) {
        nb[4] = ((uint32_t)(x2 + 1 == g.L2 ? 0 : x2 + 1) * L1 + (uint32_t)x1) * L0 + (uint32_t)x0;
        nb[5] = ((uint32_t)(x2 == 0 ? g.L2 - 1 : x2 - 1) * L1 + (uint32_t)x1) * L0 + (uint32_t)x0;
    }
    return base + (uint32_t)x0;
}

// Offsets in the other colour's plane of the neighbours of colour-c site k,
// in the same column order: nbo[j] = nb[j] >> 1 for cb_site's nb.  Site
// (x0, r) sits at r M + (x0 >> 1) of its plane, so the +-x neighbours are in
// row r at column m or m -+ 1 (by the parity of x0, with the wrap), and the
// +-y / +-z neighbours at k +- M / k +- L1 M (with the wraps).  Needs
// k < 2^31 and one FastDiv per axis beyond x, no multiplies.
template <int Z>
__host__ __device__ __forceinline__ void cb_other(const CbGeo& g, uint32_t k, int c, uint32_t (&nbo)[Z])
{
    const uint32_t M = (uint32_t)g.M;
    const uint32_t r = fdiv(k, g.divM);
    const uint32_t m = k - r * M;
    uint32_t x1 = 0, x2 = 0;
    if (Z >= 4) {
        x2 = Z == 6 ? fdiv(r, g.divL1) : 0;
        x1 = r - x2 * (uint32_t)g.L1;
    }
    const bool odd = ((x1 + x2 + (uint32_t)c) & 1) != 0;  // x0 = 2m + 1
    nbo[0] = odd ? (m == M - 1 ? k - (M - 1) : k + 1) : k;
    nbo[1] = odd ? k : (m == 0 ? k + (M - 1) : k - 1);
    if (Z >= 4) {
        const uint32_t wrap1 = ((uint32_t)g.L1 - 1) * M;
        nbo[2] = x1 + 1 == (uint32_t)g.L1 ? k - wrap1 : k + M;
        nbo[3] = x1 == 0 ? k + wrap1 : k - M;
    }
    if (Z == 6) {
        const uint32_t plane = (uint32_t)g.L1 * M, wrap2 = ((uint32_t)g.L2 - 1) * plane;
        nbo[4] = x2 + 1 == (uint32_t)g.L2 ? k - wrap2 : k + plane;
        nbo[5] = x2 == 0 ? k + wrap2 : k - plane;
    }
}

inline int cb_geo_from(int ndim, const int32_t* dims, CbGeo* g)
{
    if (ndim < 1 || ndim > 3) return -1;
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) {
        if (dims[a] < 2 || dims[a] % 2) return -2;
        N *= dims[a];
    }
    g->d = ndim;
    g->L0 = dims[0];
    g->L1 = ndim >= 2 ? dims[1] : 1;
    g->L2 = ndim >= 3 ? dims[2] : 1;
    g->M = g->L0 / 2;
    g->N = N;
    g->half = N / 2;
    cb_finish(*g);
    return 0;
}

}  // namespace cm
