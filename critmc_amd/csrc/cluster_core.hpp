// Device helpers shared by the cluster updates of every model: site geometry
// from a linear index, a lock-free union-find whose roots are the smallest
// site of each component, and the per-update Philox words of tag TAG_CLUSTER.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cb_geom.hpp"
#include "philox.hpp"

namespace cm {

constexpr uint32_t TAG_CLUSTER = 9;

struct ClGeo {
    int d;
    uint32_t L[3];
    uint32_t N;
    FastDiv div0, div1;
    // null: periodic hypercubic neighbours from L; else a neighbour table
    // [N][zt] in pair form (column 2a = +e_a, 2a+1 = -e_a), any lattice
    const int32_t* nbr;
    int zt;
};

// Bond axes: bond b = a N + i joins i and its +e_a neighbour, a < cl_axes.
__host__ __device__ __forceinline__ int cl_axes(const ClGeo& g) { return g.nbr ? g.zt / 2 : g.d; }

// Neighbour column c (2a = +e_a, 2a+1 = -e_a) of site i.
__device__ __forceinline__ uint32_t cl_nbr_col(const ClGeo& g, uint32_t i, int c);

// Coordinates of site i (x fastest).
__device__ __forceinline__ void cl_coords(const ClGeo& g, uint32_t i, uint32_t (&x)[3])
{
    const uint32_t q0 = fdiv(i, g.div0);
    x[0] = i - q0 * g.L[0];
    x[1] = 0;
    x[2] = 0;
    if (g.d >= 2) {
        const uint32_t q1 = fdiv(q0, g.div1);
        x[1] = q0 - q1 * g.L[1];
        x[2] = q1;
    }
}

__device__ __forceinline__ uint32_t cl_stride(const ClGeo& g, int a)
{
    return a == 0 ? 1u : (a == 1 ? g.L[0] : g.L[0] * g.L[1]);
}

// Neighbour of site i with coordinates x along axis a, direction +1 / -1.
__device__ __forceinline__ uint32_t cl_nbr(const ClGeo& g, uint32_t i, const uint32_t (&x)[3],
                                           int a, bool fwd)
{
    const uint32_t st = cl_stride(g, a), L = g.L[a];
    if (fwd) return x[a] + 1 == L ? i - (L - 1) * st : i + st;
    return x[a] == 0 ? i + (L - 1) * st : i - st;
}

__device__ __forceinline__ uint32_t cl_nbr_col(const ClGeo& g, uint32_t i, int c)
{
    if (g.nbr) return (uint32_t)g.nbr[(int64_t)i * g.zt + c];
    uint32_t x[3];
    cl_coords(g, i, x);
    return cl_nbr(g, i, x, c >> 1, (c & 1) == 0);
}

__device__ __forceinline__ uint32_t ld_relaxed(const int32_t* p)
{
    return (uint32_t)__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed(int32_t* p, uint32_t v)
{
    __hip_atomic_store(p, (int32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; parents only ever point to smaller sites.
__device__ __forceinline__ uint32_t cl_find(int32_t* parent, uint32_t x)
{
    while (true) {
        const uint32_t p = ld_relaxed(parent + x);
        if (p == x) return x;
        const uint32_t gp = ld_relaxed(parent + p);
        if (gp != p) st_relaxed(parent + x, gp);
        x = gp;
    }
}

// Root of x without writes: while parents are being flattened each site's
// owner is the only writer of its entry, so every write stores a root.
__device__ __forceinline__ uint32_t cl_find_ro(const int32_t* parent, uint32_t x)
{
    while (true) {
        const uint32_t p = ld_relaxed(parent + x);
        if (p == x) return x;
        x = p;
    }
}

__device__ __forceinline__ void cl_unite(int32_t* parent, uint32_t i, uint32_t j)
{
    while (true) {
        i = cl_find(parent, i);
        j = cl_find(parent, j);
        if (i == j) return;
        if (i > j) {
            const uint32_t t = i;
            i = j;
            j = t;
        }
        // hook the larger root below the smaller; fails if j stopped being a root
        if (atomicCAS(parent + j, (int32_t)j, (int32_t)i) == (int32_t)j) return;
    }
}

__device__ __forceinline__ uint32_t cl_word(uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep,
                                            uint32_t field, uint32_t index)
{
    const u32x4 r = philox4x32_10(
        u32x4{index >> 2, (uint32_t)sweep, ctr_z(sweep, field, TAG_CLUSTER), sid}, k0, k1);
    const uint32_t w = index & 3;
    return w == 0 ? r.x : (w == 1 ? r.y : (w == 2 ? r.z : r.w));
}

}  // namespace cm
