// Flat-histogram random walks in energy space, batched as many walkers:
// Wang-Landau (with the 1/t refinement) and multicanonical production
// (critmc/kernels/flathist.py:22-132, critmc/flathist.py:174-545).
//
// Energy levels (kernels/flathist.py:3-8): Ising idx = (E_b + B) / 4 (every flip
// changes E_b by a multiple of 4), Potts idx = E + B, B = number of bonds.
// A move that would leave the walker's window [lo, hi] is rejected and the
// current level is counted as for a rejected move (Schulz et al. 2003).
// Acceptance is min(1, exp(ln_w_new - ln_w_old)) with ln_w = -ln g (Wang-
// Landau) or the given weights (multicanonical): accepted iff
// ln_w_new >= ln_w_old or u < det_exp(ln_w_new - ln_w_old), so the host
// replay takes the same decisions.
//
// One lane per walker: the walk is sequential, so the GPU's parallelism is
// across walkers.  Walker r's state: int8 spins [N], level idx, ln g (or ln w)
// and visit histogram [levels], window, move counter.  Move t (the walker's
// own count) draws Philox block 0 at counter t, tag TAG_FLAT, stream sid:
// word 0 -> site floor(x N / 2^32), word 1 -> Potts proposal
// randbelow(q - 1) (+1 if >= old), word 2 -> u = x 2^-32.
#include <algorithm>
#include <cmath>

#include "cluster_core.hpp"
#include "common.hpp"
#include "det_exp.hpp"
#include "philox.hpp"

namespace cm {

constexpr uint32_t TAG_FLAT = 10;

struct FlatArgs {
    int8_t* spins;          // [R][N]
    int32_t* idx;           // [R] current level
    double* ln_g;           // [R][S] (Wang-Landau: updated) or ln_w (multicanonical: read)
    int64_t* hist;          // [R][S]
    const int32_t* lo;      // [R]
    const int32_t* hi;      // [R]
    const uint32_t* keys;   // [R][2]
    const uint32_t* sids;   // [R]
    uint64_t* moves;        // [R] moves done so far (the RNG counter)
    ClGeo g;
    int R, q, S;
};

// Level change of flipping / recolouring site i; returns the proposed state.
__device__ __forceinline__ int flat_delta(const FlatArgs& a, const int8_t* S, uint32_t i,
                                          uint32_t xprop, int* d_idx)
{
    uint32_t x[3];
    cl_coords(a.g, i, x);
    const int z = 2 * a.g.d;
    const int si = S[i];
    if (a.q == 0) {
        int field = 0;
        for (int c = 0; c < z; ++c) field += S[cl_nbr(a.g, i, x, c >> 1, (c & 1) == 0)];
        *d_idx = (si * field) / 2;  // dE_b = 2 s_i h, divided by 4
        return -si;
    }
    int nw = (int)(((uint64_t)xprop * (uint32_t)(a.q - 1)) >> 32);
    if (nw >= si) ++nw;
    int n_old = 0, n_new = 0;
    for (int c = 0; c < z; ++c) {
        const int v = S[cl_nbr(a.g, i, x, c >> 1, (c & 1) == 0)];
        n_old += v == si;
        n_new += v == nw;
    }
    *d_idx = n_old - n_new;
    return nw;
}

__device__ __forceinline__ bool flat_accept(double ln_w_old, double ln_w_new, uint32_t xu)
{
    return ln_w_new >= ln_w_old || (double)xu * 0x1p-32 < det_exp(ln_w_new - ln_w_old);
}

// Wang-Landau: n_moves moves per walker.  mode_1t[r] != 0 uses
// ln f = n_levels[r] / (t + 1) with t the walker's move count (Belardinelli-Pereyra).
__global__ __launch_bounds__(64) void wl_walk(FlatArgs a, const double* __restrict__ ln_f_in,
                                              const int32_t* __restrict__ mode_1t,
                                              const double* __restrict__ n_levels,
                                              const int32_t* __restrict__ active, int64_t n_moves)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R || (active && active[r] == 0)) return;
    const uint32_t N = a.g.N;
    int8_t* S = a.spins + (int64_t)r * N;
    double* lg = a.ln_g + (int64_t)r * a.S;
    int64_t* H = a.hist + (int64_t)r * a.S;
    const uint32_t k0 = a.keys[2 * r], k1 = a.keys[2 * r + 1], sid = a.sids[r];
    const int lo = a.lo[r], hi = a.hi[r];
    const bool one_t = mode_1t[r] != 0;
    const double nl = n_levels[r];
    double ln_f = ln_f_in[r];
    int idx = a.idx[r];
    if (lo < 0 || hi >= a.S || idx < 0 || idx >= a.S) return;  // outside the level table
    uint64_t t = a.moves[r];
    for (int64_t m = 0; m < n_moves; ++m, ++t) {
        if (one_t) ln_f = nl / ((double)t + 1.0);
        const u32x4 w = philox4x32_10(u32x4{0u, (uint32_t)t, ctr_z(t, 0, TAG_FLAT), sid}, k0, k1);
        const uint32_t i = (uint32_t)(((uint64_t)w.x * N) >> 32);
        int d;
        const int nw = flat_delta(a, S, i, w.y, &d);
        const int nidx = idx + d;
        if (lo <= nidx && nidx <= hi && flat_accept(-lg[idx], -lg[nidx], w.z)) {
            S[i] = (int8_t)nw;
            idx = nidx;
        }
        lg[idx] += ln_f;
        H[idx] += 1;
    }
    a.idx[r] = idx;
    a.moves[r] = t;
}

// Multicanonical walk with fixed weights ln_w: n_meas measurements of
// moves_per_meas moves; records the level and M (Ising) or sum_k n_k^2 (Potts).
__global__ __launch_bounds__(64) void muca_walk(FlatArgs a, int32_t* __restrict__ order,
                                                int64_t n_meas, int64_t moves_per_meas,
                                                int32_t* __restrict__ idx_out,
                                                int64_t* __restrict__ ord_out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    const uint32_t N = a.g.N;
    int8_t* S = a.spins + (int64_t)r * N;
    const double* lw = a.ln_g + (int64_t)r * a.S;
    int64_t* H = a.hist + (int64_t)r * a.S;
    const uint32_t k0 = a.keys[2 * r], k1 = a.keys[2 * r + 1], sid = a.sids[r];
    const int lo = a.lo[r], hi = a.hi[r];
    int32_t* cnt = order + (int64_t)r * (a.q == 0 ? 1 : a.q);  // M, or counts [q]
    int idx = a.idx[r];
    if (lo < 0 || hi >= a.S || idx < 0 || idx >= a.S) return;  // outside the level table
    uint64_t t = a.moves[r];
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        for (int64_t m = 0; m < moves_per_meas; ++m, ++t) {
            const u32x4 w = philox4x32_10(u32x4{0u, (uint32_t)t, ctr_z(t, 0, TAG_FLAT), sid}, k0, k1);
            const uint32_t i = (uint32_t)(((uint64_t)w.x * N) >> 32);
            int d;
            const int old = S[i];
            const int nw = flat_delta(a, S, i, w.y, &d);
            const int nidx = idx + d;
            if (lo <= nidx && nidx <= hi && flat_accept(lw[idx], lw[nidx], w.z)) {
                S[i] = (int8_t)nw;
                idx = nidx;
                if (a.q == 0) {
                    cnt[0] -= 2 * old;
                } else {
                    cnt[old] -= 1;
                    cnt[nw] += 1;
                }
            }
            H[idx] += 1;
        }
        const int64_t o = (int64_t)r * n_meas + meas;
        idx_out[o] = idx;
        if (a.q == 0) {
            ord_out[o] = cnt[0];
        } else {
            int64_t sq = 0;
            for (int k = 0; k < a.q; ++k) sq += (int64_t)cnt[k] * cnt[k];
            ord_out[o] = sq;
        }
    }
    a.idx[r] = idx;
    a.moves[r] = t;
}

inline int flat_args(FlatArgs& a, void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                     int32_t q, int32_t* idx, double* ln_g, int64_t* hist, int32_t n_levels,
                     const int32_t* lo, const int32_t* hi, const uint32_t* keys,
                     const uint32_t* sids, uint64_t* moves)
{
    CM_REQUIRE(ndim >= 1 && ndim <= 3, "lattices of 1..3 dimensions");
    int64_t N = 1;
    for (int k = 0; k < ndim; ++k) {
        CM_REQUIRE(dims[k] >= 2, "extents must be >= 2");
        N *= dims[k];
    }
    CM_REQUIRE(N < ((int64_t)1 << 31), "lattice too large");
    CM_REQUIRE(n_rep >= 1, "need at least one walker");
    CM_REQUIRE(q == 0 || (q >= 2 && q <= 127), "q must be 0 (Ising) or 2..127 (Potts)");
    const int64_t bonds = N * ndim;
    const int64_t levels = q == 0 ? bonds / 2 + 1 : bonds + 1;
    CM_REQUIRE(n_levels == levels, "expected %lld energy levels, got %d", (long long)levels, n_levels);
    a.spins = (int8_t*)spins;
    a.idx = idx;
    a.ln_g = ln_g;
    a.hist = hist;
    a.lo = lo;
    a.hi = hi;
    a.keys = keys;
    a.sids = sids;
    a.moves = moves;
    a.g.d = ndim;
    for (int k = 0; k < 3; ++k) a.g.L[k] = k < ndim ? (uint32_t)dims[k] : 1u;
    a.g.N = (uint32_t)N;
    a.g.div0 = make_fastdiv(a.g.L[0]);
    a.g.div1 = make_fastdiv(a.g.L[1]);
    a.R = n_rep;
    a.q = q;
    a.S = n_levels;
    return CM_OK;
}

}  // namespace cm

using namespace cm;

extern "C" {

int cm_wl_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
              int32_t* idx, double* ln_g, int64_t* hist, int32_t n_levels, const int32_t* lo,
              const int32_t* hi, const uint32_t* keys, const uint32_t* stream_ids, uint64_t* moves,
              const double* ln_f, const int32_t* mode_1t, const double* n_levels_1t,
              const int32_t* active, int64_t n_moves, void* stream)
{
    FlatArgs a{};
    if (int rc = flat_args(a, spins, ndim, dims, n_rep, q, idx, ln_g, hist, n_levels, lo, hi, keys,
                           stream_ids, moves))
        return rc;
    CM_REQUIRE(n_moves >= 0, "n_moves must not be negative");
    if (n_moves == 0) return CM_OK;
    hipLaunchKernelGGL(wl_walk, dim3((n_rep + 63) / 64), dim3(64), 0, (hipStream_t)stream, a, ln_f,
                       mode_1t, n_levels_1t, active, n_moves);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_muca_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
                int32_t* idx, int32_t* order, const double* ln_w, int64_t* hist, int32_t n_levels,
                const int32_t* lo, const int32_t* hi, const uint32_t* keys,
                const uint32_t* stream_ids, uint64_t* moves, int64_t n_meas,
                int64_t moves_per_meas, int32_t* idx_out, int64_t* ord_out, void* stream)
{
    FlatArgs a{};
    if (int rc = flat_args(a, spins, ndim, dims, n_rep, q, idx, (double*)ln_w, hist, n_levels, lo,
                           hi, keys, stream_ids, moves))
        return rc;
    CM_REQUIRE(n_meas >= 0 && moves_per_meas >= 1, "bad measurement counts");
    if (n_meas == 0) return CM_OK;
    hipLaunchKernelGGL(muca_walk, dim3((n_rep + 63) / 64), dim3(64), 0, (hipStream_t)stream, a,
                       order, n_meas, moves_per_meas, idx_out, ord_out);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // extern "C"
