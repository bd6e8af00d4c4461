// Checkerboard q-state Potts sweeps (int8 states 0..q-1), Metropolis and heat bath.
//
// Reference semantics (critmc/kernels/potts.py):
//   Metropolis :44-71  new = randbelow(q-1), +1 if >= old; delta = n_old - n_new;
//                      accept iff delta <= 0 or u < exp(-beta delta) (table :194-197)
//   heat bath  :74-103 w[n] = exp(beta n) (:198); r = u * sum_k w[n_k]; new = first k
//                      with r < cumulative sum (k ascending), default q-1
//   run        :170-235 per measurement E = -#satisfied bonds, counts[q], accepted
//
// Up to 96 KB of states per replica (L <= 256 in 2D): one workgroup owns one
// replica at a time (grid-stride over replicas) and runs the whole call's
// sweeps with the states in LDS.  Larger lattices, and batches too small to
// fill the CUs that way, take the wide form: one launch per half sweep over
// all replicas, states in global memory.  Both forms draw the same
// uniforms for the same site and sweep, so they give identical chains.
//
// Uniforms: the colour-c sites are enumerated k = 0 .. N/2-1 (row-major, see
// cb_geom.hpp), tag TAG_POTTS.  Metropolis: site k uses words 2(k&1) and
// 2(k&1)+1 of the Philox block k >> 1, the proposal (randbelow(n) =
// floor(x n / 2^32)) and the acceptance u.  Heat bath: site k uses word k&3 of
// block k >> 2.  A lane updates a quad of colour sites 4u .. 4u+3.  The
// heat-bath arithmetic is FP64 in the reference's order, compiled without
// contraction, so it is bit-reproducible on the host.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cb_geom.hpp"
#include "common.hpp"
#include "ising_core.hpp"
#include "philox.hpp"

namespace cm {

struct alignas(16) PottsTab {
    uint64_t T[8];      // Metropolis: accept iff x < T[delta] (delta >= 1); T = ceil(exp(-beta d) 2^32)
    double w[16];       // heat bath: exp(beta n) for n <= z; 0 elsewhere (w[15]: padding states)
    double w2[64][2];   // w2[a + 8 b] = {w[a], w[b]}: two consecutive 3-bit histogram fields at once
};

constexpr int POTTS_QMAX = 127;
constexpr int64_t POTTS_LDS_STATE = 96 * 1024;
// Auto choice of the wide form, from measured crossovers on MI355X (q = 10):
// one workgroup per replica needs enough replicas to fill the CUs, and at 64
// KB of LDS per replica (L = 256) only two fit on a CU.
static bool potts_prefer_wide(int64_t N, int n_rep)
{
    return N > POTTS_LDS_STATE || N >= 65536 || (N >= 16384 && n_rep < 1024) ||
           (N >= 4096 && n_rep < 256);
}

// One site update; returns the new state (== old when the site keeps it) and
// adds the energy change and the acceptance to de / acc.  tl is the replica's
// table staged in LDS.
//
// QB > 0 (q <= QB <= 16): the neighbour states are folded into a histogram of
// 4-bit fields, h = pad + sum_j 16^{s_j}, so n_k = (h >> 4k) & 15 is one
// bit-field extract; pad sets the fields of the states q .. QB-1 to 15, whose
// weight w[15] is 0, so all QB terms can be added without a branch.  The heat
// bath adds w[n_k] for k = 0 .. q-1 in that order, as the reference does,
// keeping every prefix sum; since the prefix sums never decrease, the first k
// with r < cum_k equals the number of k <= q-2 with !(r < cum_k), and the
// default q-1 falls out of the same count.  The padded terms do not change
// the count: their prefix sums equal the total, and r = u * total < total.
// QB == 0: any q <= 127, counts recomputed per state.
template <int Z, int QB>
__device__ __forceinline__ int potts_update(int8_t* __restrict__ S, uint32_t i,
                                            const uint32_t (&nbi)[Z], int q, uint64_t pad,
                                            int code, uint32_t xa, uint32_t xb,
                                            const PottsTab* __restrict__ tl, int& de, int& acc,
                                            int& old_out)
{
    const int old = S[i];
    old_out = old;
    int nb[Z];
#pragma unroll
    for (int j = 0; j < Z; ++j) nb[j] = S[nbi[j]];
    int nw, c_old, c_new;
    if constexpr (QB > 0) {
        uint64_t h = pad;
#pragma unroll
        for (int j = 0; j < Z; ++j) h += 1ull << (4 * nb[j]);
        c_old = (int)(h >> (4 * old)) & 15;
        if (code == CODE_METROPOLIS) {
            nw = (int)(((uint64_t)xa * (uint32_t)(q - 1)) >> 32);
            if (nw >= old) ++nw;
            c_new = (int)(h >> (4 * nw)) & 15;
            const int delta = c_old - c_new;
            if (!(delta <= 0 || (uint64_t)xb < tl->T[delta])) return old;
        } else {
            double cum[QB];
            double total = 0.0;
#pragma unroll
            for (int k = 0; k < QB; ++k) {
                total = __dadd_rn(total, tl->w[(uint32_t)(h >> (4 * k)) & 15u]);
                cum[k] = total;
            }
            const double r = __dmul_rn((double)xa * 0x1p-32, total);
            nw = 0;
#pragma unroll
            for (int k = 0; k < QB - 1; ++k) nw += !(r < cum[k]);
            if (nw == old) return old;
            c_new = (int)(h >> (4 * nw)) & 15;
        }
    } else {
        if (code == CODE_METROPOLIS) {
            nw = (int)(((uint64_t)xa * (uint32_t)(q - 1)) >> 32);
            if (nw >= old) ++nw;
            c_old = 0;
            c_new = 0;
#pragma unroll
            for (int j = 0; j < Z; ++j) {
                c_old += nb[j] == old;
                c_new += nb[j] == nw;
            }
            const int delta = c_old - c_new;
            if (!(delta <= 0 || (uint64_t)xb < tl->T[delta])) return old;
        } else {
            double total = 0.0;
            for (int k = 0; k < q; ++k) {
                int c = 0;
#pragma unroll
                for (int j = 0; j < Z; ++j) c += nb[j] == k;
                total = __dadd_rn(total, tl->w[c]);
            }
            const double r = __dmul_rn((double)xa * 0x1p-32, total);
            nw = q - 1;
            double cum = 0.0;
            for (int k = 0; k < q; ++k) {
                int c = 0;
#pragma unroll
                for (int j = 0; j < Z; ++j) c += nb[j] == k;
                cum = __dadd_rn(cum, tl->w[c]);
                if (r < cum) {
                    nw = k;
                    break;
                }
            }
            if (nw == old) return old;
            c_old = 0;
            c_new = 0;
#pragma unroll
            for (int j = 0; j < Z; ++j) {
                c_old += nb[j] == old;
                c_new += nb[j] == nw;
            }
        }
    }
    de += c_old - c_new;
    ++acc;
    S[i] = (int8_t)nw;
    return nw;
}

// Fields q .. QB-1 of the padded histogram set to 15.
template <int QB>
__device__ __forceinline__ uint64_t potts_pad(int q)
{
    uint64_t pad = 0;
    for (int k = q; k < QB; ++k) pad |= 15ull << (4 * k);
    return pad;
}

// Copy one replica's table into LDS.
__device__ __forceinline__ void stage_tab(PottsTab* __restrict__ dst, const PottsTab* __restrict__ src)
{
    for (uint32_t t = threadIdx.x; t < sizeof(PottsTab) / 16; t += blockDim.x)
        ((uint4*)dst)[t] = ((const uint4*)src)[t];
}

#ifndef POTTS_HB_SITES
#define POTTS_HB_SITES 2
#endif

// Heat-bath choices of NS sites with 3-bit histograms h: the weights of fields
// 2i and 2i+1 come in one 16-byte read, total = w[n_0] + w[n_1] + ... in
// order k = 0 .. QB-1 keeping every prefix sum, r = u * total, and the new
// state is the number of k <= QB-2 with !(r < cum_k) (potts_update's rule).
template <int QB, int NS>
__device__ __forceinline__ void potts_hb_decide(const double2* __restrict__ w2, const uint32_t* h,
                                                const uint32_t* x, int* nw)
{
    static_assert(QB % 2 == 0, "fields are read in pairs");
    double cum[NS][QB], total[NS];
#pragma unroll
    for (int k = 0; k < QB; k += 2)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const double2 wp = w2[(h[j] >> (3 * k)) & 63u];
            total[j] = k == 0 ? wp.x : __dadd_rn(total[j], wp.x);  // 0 + w[n_0]: exact
            cum[j][k] = total[j];
            total[j] = __dadd_rn(total[j], wp.y);
            cum[j][k + 1] = total[j];
        }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const double rr = __dmul_rn((double)x[j] * 0x1p-32, total[j]);
        int n = 0;
#pragma unroll
        for (int k = 0; k < QB - 1; ++k) n += !(rr < cum[j][k]);
        nw[j] = n;
    }
}

// Square lattices with L0 % 8 == 0 and q <= QB <= 10: the quad's four colour
// sites x0 = x8 + 2j + p (j = 0..3) of row x1 lie in the aligned 8-byte span
// [x8, x8 + 8) of their row, so the quad reads three 8-byte words (its row and
// the rows above and below) and the one x-neighbour just outside the span.
// Its row is realigned into a window whose byte 2j+1 is site j, 2j its -x and
// 2j+2 its +x neighbour.  Bytes are scaled to 3 s (states <= 9: no carry
// between bytes), and the histogram has 3-bit fields, h = pad + sum_j 8^{s_j}
// (fields q .. QB-1 set to 7, whose weight w[7] is 0).  The decisions, and
// the FP64 sums behind them, are those of potts_update.
template <int QB>
__device__ __forceinline__ void potts_quad2d(int8_t* __restrict__ S, const CbGeo& g, uint32_t u,
                                             int c, int q, uint32_t pad3, int code, uint64_t sweep,
                                             uint32_t zc, uint32_t k0, uint32_t k1, uint32_t sid,
                                             const PottsTab* __restrict__ tl, int& de, int& acc,
                                             int* __restrict__ counts, const double2* __restrict__ w2)
{
    static_assert(QB > 0 && QB <= 10, "3-bit histogram fields for at most 10 states");
    uint32_t xa[4], xb[4];
    if (code == CODE_METROPOLIS) {
        const u32x4 r0 = philox4x32_10(u32x4{2 * u, (uint32_t)sweep, zc, sid}, k0, k1);
        const u32x4 r1 = philox4x32_10(u32x4{2 * u + 1, (uint32_t)sweep, zc, sid}, k0, k1);
        xa[0] = r0.x; xb[0] = r0.y; xa[1] = r0.z; xb[1] = r0.w;
        xa[2] = r1.x; xb[2] = r1.y; xa[3] = r1.z; xb[3] = r1.w;
    } else {
        const u32x4 r0 = philox4x32_10(u32x4{u, (uint32_t)sweep, zc, sid}, k0, k1);
        xa[0] = r0.x; xa[1] = r0.y; xa[2] = r0.z; xa[3] = r0.w;
        xb[0] = xb[1] = xb[2] = xb[3] = 0;
    }
    const uint32_t M = (uint32_t)g.M, L0 = (uint32_t)g.L0, L1 = (uint32_t)g.L1;
    const uint32_t k4 = 4u * u;
    const uint32_t r = fdiv(k4, g.divM);  // the row, x1
    const uint32_t x8 = 2u * (k4 - r * M);
    const uint32_t p = (r + (uint32_t)c) & 1u;
    const uint32_t row = r * L0;
    const uint32_t rup = (r + 1 == L1 ? 0u : r + 1) * L0, rdn = (r == 0 ? L1 - 1 : r - 1) * L0;
    const uint2* const S2 = (const uint2*)S;  // 8-byte words: 32-bit indices off one base
    const uint2 own = S2[(row + x8) >> 3];
    const uint2 up = S2[(rup + x8) >> 3];
    const uint2 dn = S2[(rdn + x8) >> 3];
    const uint32_t xe = p ? (x8 + 8 == L0 ? 0u : x8 + 8) : (x8 == 0 ? L0 - 1 : x8 - 1);
    const uint32_t ex = (uint32_t)(uint8_t)S[row + xe];
    const uint32_t A = p ? own.x : (own.x << 8) | ex;
    const uint32_t B = p ? own.y : (own.y << 8) | (own.x >> 24);
    const uint32_t C = p ? ex : own.y >> 24;
    const uint32_t A3 = A + (A << 1), B3 = B + (B << 1), C3 = C + (C << 1);
    const uint32_t U3[2] = {up.x + (up.x << 1), up.y + (up.y << 1)};
    const uint32_t D3[2] = {dn.x + (dn.x << 1), dn.y + (dn.y << 1)};
    const uint32_t sh = 8u * p;
    uint32_t old[4], h[4];
    int c_old[4], nw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t wo = j < 2 ? A : B, wo3 = j < 2 ? A3 : B3;
        const uint32_t bo = 8u * (uint32_t)((2 * j + 1) & 3);
        old[j] = (wo >> bo) & 0xFFu;
        const uint32_t old3 = (wo3 >> bo) & 0xFFu;
        const uint32_t mx3 = (wo3 >> (bo - 8)) & 0xFFu;
        const uint32_t px3 = j == 3 ? C3 : ((j == 1 ? B3 : wo3) >> ((bo + 8) & 31)) & 0xFFu;
        const uint32_t yb = 16u * (uint32_t)(j & 1) + sh;
        const uint32_t uy3 = __builtin_amdgcn_ubfe(U3[j >> 1], yb, 8);
        const uint32_t dy3 = __builtin_amdgcn_ubfe(D3[j >> 1], yb, 8);
        h[j] = pad3 + (1u << px3) + (1u << mx3) + (1u << uy3) + (1u << dy3);
        c_old[j] = (int)((h[j] >> old3) & 7u);
    }
    if (code == CODE_METROPOLIS) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int n = (int)(((uint64_t)xa[j] * (uint32_t)(q - 1)) >> 32);
            if (n >= (int)old[j]) ++n;
            const int delta = c_old[j] - (int)((h[j] >> (3 * n)) & 7u);
            nw[j] = delta <= 0 || (uint64_t)xb[j] < tl->T[delta] ? n : (int)old[j];
        }
    } else {
        // The sites' FP64 chains side by side (POTTS_HB_SITES at a time), so
        // their dependent additions and table reads overlap.
#pragma unroll
        for (int j0 = 0; j0 < 4; j0 += POTTS_HB_SITES)
            potts_hb_decide<QB, POTTS_HB_SITES>(w2, h + j0, xa + j0, nw + j0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (nw[j] == (int)old[j]) continue;
        de += c_old[j] - (int)((h[j] >> (3 * nw[j])) & 7u);
        ++acc;
        S[row + x8 + 2 * j + p] = (int8_t)nw[j];
        atomicAdd(&counts[old[j]], -1);
        atomicAdd(&counts[nw[j]], 1);
    }
}

__device__ __forceinline__ uint32_t potts_pad3(int q, int qb)
{
    uint32_t pad = 0;
    for (int k = q; k < qb; ++k) pad |= 7u << (3 * k);
    return pad;
}

// The quad of colour-c sites 4u .. 4u+3 (those below `half`): Metropolis draws
// blocks 2u and 2u+1 (two words per site), the heat bath block u (one word
// per site).  Occupation changes go to the LDS counts with atomics.
template <int Z, int QB>
__device__ __forceinline__ void potts_quad(int8_t* __restrict__ S, const CbGeo& g, uint32_t u,
                                           int c, int q, uint64_t pad, int code, uint64_t sweep,
                                           uint32_t zc, uint32_t k0, uint32_t k1, uint32_t sid,
                                           const PottsTab* __restrict__ tl, int& de, int& acc,
                                           int* __restrict__ counts)
{
    if constexpr (Z == 4 && QB > 0 && QB <= 10) {
        if ((g.L0 & 7) == 0) {
            potts_quad2d<QB>(S, g, u, c, q, potts_pad3(q, QB), code, sweep, zc, k0, k1, sid, tl, de,
                             acc, counts, (const double2*)tl->w2);
            return;
        }
    }
    uint32_t xa[4], xb[4];
    if (code == CODE_METROPOLIS) {
        const u32x4 r0 = philox4x32_10(u32x4{2 * u, (uint32_t)sweep, zc, sid}, k0, k1);
        const u32x4 r1 = philox4x32_10(u32x4{2 * u + 1, (uint32_t)sweep, zc, sid}, k0, k1);
        xa[0] = r0.x; xb[0] = r0.y; xa[1] = r0.z; xb[1] = r0.w;
        xa[2] = r1.x; xb[2] = r1.y; xa[3] = r1.z; xb[3] = r1.w;
    } else {
        const u32x4 r0 = philox4x32_10(u32x4{u, (uint32_t)sweep, zc, sid}, k0, k1);
        xa[0] = r0.x; xa[1] = r0.y; xa[2] = r0.z; xa[3] = r0.w;
        xb[0] = xb[1] = xb[2] = xb[3] = 0;
    }
    const uint32_t half = (uint32_t)g.half;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t k = 4 * u + j;
        if (k < half) {
            uint32_t nbi[Z];
            const uint32_t i = cb_site32<Z>(g, k, c, nbi);
            int old;
            const int nw = potts_update<Z, QB>(S, i, nbi, q, pad, code, xa[j], xb[j], tl, de, acc, old);
            if (nw != old) {
                atomicAdd(&counts[old], -1);
                atomicAdd(&counts[nw], 1);
            }
        }
    }
}

template <int Z, int QB>
__global__ __launch_bounds__(256) void potts_batched(
    int8_t* __restrict__ spins, CbGeo g, int q, int R, const PottsTab* __restrict__ tabs,
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ sids,
    uint64_t* __restrict__ sweeps, Sched sc, int64_t n_meas, int32_t* __restrict__ e_run,
    int32_t* __restrict__ counts_run, int32_t* __restrict__ e_out, int32_t* __restrict__ acc_out,
    int32_t* __restrict__ counts_out, int use_lds, const int64_t* __restrict__ n_meas_rep)
{
    extern __shared__ int4 lds4[];
    PottsTab* tl = (PottsTab*)lds4;
    int* counts = (int*)(tl + 1);      // [128]
    int* red = counts + 128;           // [16][2]
    int8_t* lstate = (int8_t*)(red + 32);
    const int nwaves = (blockDim.x + 63) >> 6;
    const uint32_t nquads = (uint32_t)((g.half + 3) >> 2);
    const uint64_t pad = potts_pad<QB>(q);

    for (int rep = blockIdx.x; rep < R; rep += gridDim.x) {
        int8_t* gs = spins + (int64_t)rep * g.N;
        int8_t* S = use_lds ? lstate : gs;
        if (use_lds) {
            if ((g.N & 15) == 0) {
                const int4* src = (const int4*)gs;
                int4* dst = (int4*)lstate;
                for (int64_t t = threadIdx.x; t < g.N / 16; t += blockDim.x) dst[t] = src[t];
            } else {
                for (int64_t t = threadIdx.x; t < g.N; t += blockDim.x) lstate[t] = gs[t];
            }
        }
        stage_tab(tl, tabs + rep);
        for (int k = threadIdx.x; k < q; k += blockDim.x) counts[k] = counts_run[(int64_t)rep * q + k];
        __syncthreads();
        const uint32_t k0 = keys[2 * rep], k1 = keys[2 * rep + 1], sid = sids[rep];
        uint64_t sweep = sweeps[rep];
        int32_t e = e_run[rep];
        const int64_t n_rep_meas = n_meas_rep ? min(n_meas, n_meas_rep[rep]) : n_meas;
        for (int64_t meas = 0; meas < n_rep_meas; ++meas) {
            int de = 0, acc = 0;
            for (int row = 0; row < sc.n; ++row) {
                const int code = sc.code[row];
                for (int rp = 0; rp < sc.count[row]; ++rp) {
                    for (int c = 0; c < 2; ++c) {
                        const uint32_t zc = ctr_z(sweep, (uint32_t)c, TAG_POTTS);
                        for (uint32_t u = threadIdx.x; u < nquads; u += blockDim.x)
                            potts_quad<Z, QB>(S, g, u, c, q, pad, code, sweep, zc, k0, k1, sid, tl,
                                              de, acc, counts);
                        __syncthreads();
                    }
                    ++sweep;
                }
            }
            de = wave_sum(de);
            acc = wave_sum(acc);
            if ((threadIdx.x & 63) == 0) {
                red[2 * (threadIdx.x >> 6)] = de;
                red[2 * (threadIdx.x >> 6) + 1] = acc;
            }
            __syncthreads();
            int tde = 0, tac = 0;
            for (int w = 0; w < nwaves; ++w) {
                tde += red[2 * w];
                tac += red[2 * w + 1];
            }
            e += tde;
            const int64_t o = (int64_t)rep * n_meas + meas;
            if (threadIdx.x == 0) {
                e_out[o] = e;
                acc_out[o] = tac;
            }
            for (int k = threadIdx.x; k < q; k += blockDim.x) counts_out[o * q + k] = counts[k];
            __syncthreads();
        }
        if (use_lds) {
            if ((g.N & 15) == 0) {
                const int4* src = (const int4*)lstate;
                int4* dst = (int4*)gs;
                for (int64_t t = threadIdx.x; t < g.N / 16; t += blockDim.x) dst[t] = src[t];
            } else {
                for (int64_t t = threadIdx.x; t < g.N; t += blockDim.x) gs[t] = lstate[t];
            }
        }
        for (int k = threadIdx.x; k < q; k += blockDim.x) counts_run[(int64_t)rep * q + k] = counts[k];
        if (threadIdx.x == 0) {
            e_run[rep] = e;
            sweeps[rep] = sweep;
        }
        __syncthreads();
    }
}

// Wide form for lattices too large for LDS: one launch per half sweep over
// (quad chunks x replicas), states in global memory.  Energy, accepted and
// occupation deltas are reduced per workgroup and added with integer atomics,
// so every sum is exact; potts_prefix turns them into running values.
template <int Z, int QB>
__global__ __launch_bounds__(256) void potts_half_wide(
    int8_t* __restrict__ spins, CbGeo g, int q, int code, int colour, uint64_t sweep_add,
    const PottsTab* __restrict__ tabs, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ sids, const uint64_t* __restrict__ sweeps, int64_t n_meas,
    int64_t meas, const int64_t* __restrict__ n_meas_rep, int32_t* __restrict__ de_out,
    int32_t* __restrict__ acc_out, int32_t* __restrict__ dcounts_out, int qpl)
{
    __shared__ PottsTab tl;
    __shared__ int counts[128];
    __shared__ int red[2 * 4];
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x);
    const uint32_t bx = lb - (uint32_t)rep * gridDim.x;
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;  // whole block
    for (int k = threadIdx.x; k < q; k += blockDim.x) counts[k] = 0;
    stage_tab(&tl, tabs + rep);
    __syncthreads();
    int8_t* S = spins + (int64_t)rep * g.N;
    const uint64_t sweep = sweeps[rep] + sweep_add;
    const uint32_t k0 = keys[2 * rep], k1 = keys[2 * rep + 1], sid = sids[rep];
    const uint32_t nquads = (uint32_t)((g.half + 3) >> 2);
    const uint32_t zc = ctr_z(sweep, (uint32_t)colour, TAG_POTTS);
    const uint64_t pad = potts_pad<QB>(q);
    int de = 0, acc = 0;
    // qpl quads per lane; for each, the wave's lanes take consecutive quads
    for (int j = 0; j < qpl; ++j) {
        const uint32_t u = (bx * (uint32_t)qpl + (uint32_t)j) * blockDim.x + threadIdx.x;
        if (u < nquads)
            potts_quad<Z, QB>(S, g, u, colour, q, pad, code, sweep, zc, k0, k1, sid, &tl, de, acc,
                              counts);
    }
    // de in [-4Z qpl, 4Z qpl], acc in [0, 4 qpl] (qpl <= 8): bitwise ballots sum them
    const int lanes = __popcll(__ballot(1));
    const int off = 4 * Z * qpl;
    de = wave_sum_bits<10>((uint32_t)(de + off)) - off * lanes;
    acc = wave_sum_bits<6>((uint32_t)acc);
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = de;
        red[2 * (threadIdx.x >> 6) + 1] = acc;
    }
    __syncthreads();
    const int64_t o = (int64_t)rep * n_meas + meas;
    if (threadIdx.x == 0) {
        int tde = 0, tac = 0;
        for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) {
            tde += red[2 * w];
            tac += red[2 * w + 1];
        }
        if (tde) atomicAdd(&de_out[o], tde);
        if (tac) atomicAdd(&acc_out[o], tac);
    }
    for (int k = threadIdx.x; k < q; k += blockDim.x)
        if (counts[k]) atomicAdd(&dcounts_out[o * q + k], counts[k]);
}

__global__ void potts_zero(int32_t* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Deltas -> running E and counts: one thread per (replica, column), column q
// being E and column c < q the count of state c, so each running sum stays in
// a register and the replica's columns run side by side.
__global__ void potts_prefix(int32_t* __restrict__ e_out, int32_t* __restrict__ counts_out,
                             int64_t n_meas, int32_t* __restrict__ e_run,
                             int32_t* __restrict__ counts_run, int q, int R,
                             const int64_t* __restrict__ n_meas_rep, uint64_t* __restrict__ sweeps,
                             uint64_t per_meas)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)R * (q + 1)) return;
    const int rep = (int)(t / (q + 1));
    const int c = (int)(t % (q + 1));
    const int64_t n = n_meas_rep ? min(n_meas, n_meas_rep[rep]) : n_meas;
    if (c == q) {
        int32_t e = e_run[rep];
        for (int64_t k = 0; k < n; ++k) {
            const int64_t o = (int64_t)rep * n_meas + k;
            e += e_out[o];
            e_out[o] = e;
        }
        e_run[rep] = e;
        sweeps[rep] += per_meas * (uint64_t)n;
    } else {
        int32_t v = counts_run[(int64_t)rep * q + c];
        for (int64_t k = 0; k < n; ++k) {
            const int64_t o = ((int64_t)rep * n_meas + k) * q + c;
            v += counts_out[o];
            counts_out[o] = v;
        }
        counts_run[(int64_t)rep * q + c] = v;
    }
}

__global__ void potts_init_kernel(int8_t* __restrict__ s, int64_t N, int R, int q,
                                  const uint32_t* __restrict__ keys,
                                  const uint32_t* __restrict__ sids, int start)
{
    const int rep = blockIdx.y;
    const uint32_t k0 = keys[2 * rep], k1 = keys[2 * rep + 1], sid = sids[rep];
    for (int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; blk * 4 < N;
         blk += (int64_t)gridDim.x * blockDim.x) {
        u32x4 r{0, 0, 0, 0};
        if (start == 0) r = draw(k0, k1, sid, 0, 0, TAG_INIT, (uint32_t)blk);
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
        for (int k = 0; k < 4 && blk * 4 + k < N; ++k)
            s[(int64_t)rep * N + blk * 4 + k] = (int8_t)(((uint64_t)w[k] * (uint32_t)q) >> 32);
    }
}

// E = -#satisfied bonds and counts[q] per replica (critmc/kernels/potts.py:27-41).
// E and the state counts of each replica.  Workgroup b takes `chunk` of the
// replica b / nchunk's sites (colour 0, then colour 1) and adds its integer
// sums atomically onto sums zeroed by potts_measure_zero, so they do not
// depend on the order.
__global__ void potts_measure_zero(int32_t* e_run, int32_t* counts_run, int n_rep, int q)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_rep) e_run[t] = 0;
    if (t < (int64_t)n_rep * q) counts_run[t] = 0;
}

template <int Z>
__global__ __launch_bounds__(256) void potts_measure_kernel(const int8_t* __restrict__ s, CbGeo g,
                                                            int q, int64_t chunk, int64_t nchunk,
                                                            int32_t* __restrict__ e_run,
                                                            int32_t* __restrict__ counts_run)
{
    __shared__ int counts[128];
    __shared__ int red[16];
    const int64_t rep = blockIdx.x / nchunk, part = blockIdx.x % nchunk;
    const int8_t* S = s + rep * g.N;
    for (int k = threadIdx.x; k < q; k += blockDim.x) counts[k] = 0;
    __syncthreads();
    int e = 0;
    const int64_t t1 = std::min<int64_t>(2 * g.half, (part + 1) * chunk);
    for (int64_t t = part * chunk + threadIdx.x; t < t1; t += blockDim.x) {
        const int c = t >= g.half ? 1 : 0;
        const int64_t k = c ? t - g.half : t;
        {
            int64_t nbi[Z];
            const int64_t i = cb_site<Z>(g, k, c, nbi);
            const int v = S[i];
            atomicAdd(&counts[v], 1);
            // forward bonds only: columns 0, 2, 4
#pragma unroll
            for (int j = 0; j < Z; j += 2) e -= S[nbi[j]] == v;
        }
    }
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        if (t) atomicAdd(e_run + rep, t);
    }
    for (int k = threadIdx.x; k < q; k += blockDim.x)
        if (counts[k]) atomicAdd(counts_run + rep * q + k, counts[k]);
}

static uint64_t potts_threshold(double p)
{
    if (!(p < 1.0)) return 1ull << 32;
    if (!(p > 0.0)) return 0;
    return (uint64_t)std::ceil(p * 4294967296.0);
}

template <int Z, int QB>
static hipError_t launch_potts(int8_t* spins, const CbGeo& g, int q, int R, const PottsTab* tabs,
                               const uint32_t* keys, const uint32_t* sids, uint64_t* sweeps,
                               const Sched& sc, int64_t n_meas, int32_t* e_run,
                               int32_t* counts_run, int32_t* e_out, int32_t* acc_out,
                               int32_t* counts_out, const int64_t* n_meas_rep, int block,
                               int grid, hipStream_t st)
{
    const int use_lds = g.N <= POTTS_LDS_STATE ? 1 : 0;
    const size_t lds = sizeof(PottsTab) + (128 + 32) * 4 +
                       (use_lds ? (size_t)((g.N + 15) / 16) * 16 : 0);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)potts_batched<Z, QB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((potts_batched<Z, QB>), dim3(grid), dim3(block), lds, st, spins, g, q, R,
                       tabs, keys, sids, sweeps, sc, n_meas, e_run, counts_run, e_out, acc_out,
                       counts_out, use_lds, n_meas_rep);
    return hipGetLastError();
}

template <int Z>
static hipError_t launch_potts_q(int8_t* spins, const CbGeo& g, int q, int R, const PottsTab* tabs,
                                 const uint32_t* keys, const uint32_t* sids, uint64_t* sweeps,
                                 const Sched& sc, int64_t n_meas, int32_t* e_run,
                                 int32_t* counts_run, int32_t* e_out, int32_t* acc_out,
                                 int32_t* counts_out, const int64_t* n_meas_rep, int block,
                                 int grid, hipStream_t st)
{
#define CM_POTTS_B(QB) \
    launch_potts<Z, QB>(spins, g, q, R, tabs, keys, sids, sweeps, sc, n_meas, e_run, counts_run, \
                        e_out, acc_out, counts_out, n_meas_rep, block, grid, st)
    if (q <= 4) return CM_POTTS_B(4);
    if (q <= 8) return CM_POTTS_B(8);
    if (q == 10) return CM_POTTS_B(10);  // the serving mix's q: no padded states
    if (q <= 12) return CM_POTTS_B(12);
    if (q <= 16) return CM_POTTS_B(16);
    return CM_POTTS_B(0);
#undef CM_POTTS_B
}

template <int Z>
static void launch_wide(const dim3& wgrid, hipStream_t st, int8_t* spins, const CbGeo& g, int q,
                        int code, int c, uint64_t sw, const PottsTab* tabs, const uint32_t* keys,
                        const uint32_t* sids, const uint64_t* sweeps, int64_t n_meas, int64_t meas,
                        const int64_t* n_meas_rep, int32_t* e_out, int32_t* acc_out,
                        int32_t* counts_out, int qpl)
{
#define CM_POTTS_W(QB) \
    hipLaunchKernelGGL((potts_half_wide<Z, QB>), wgrid, dim3(256), 0, st, spins, g, q, code, c, \
                       sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, acc_out, \
                       counts_out, qpl)
    if (q <= 4) CM_POTTS_W(4);
    else if (q <= 8) CM_POTTS_W(8);
    else if (q == 10) CM_POTTS_W(10);
    else if (q <= 12) CM_POTTS_W(12);
    else if (q <= 16) CM_POTTS_W(16);
    else CM_POTTS_W(0);
#undef CM_POTTS_W
}

}  // namespace cm

using namespace cm;

extern "C" int64_t cm_potts_workspace_bytes(int32_t n_rep);

namespace cm {

// Tables of n_rep chains (Metropolis thresholds ceil(e^{-beta d} 2^32), heat-bath
// weights e^{beta n}; critmc/kernels/potts.py:194-200) uploaded into the workspace.
int potts_prepare(int32_t ndim, const int32_t* dims, int32_t n_rep, const double* betas,
                  void* workspace, int64_t workspace_bytes, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(n_rep >= 1, "n_rep must be positive");
    CM_REQUIRE(workspace_bytes >= cm_potts_workspace_bytes(n_rep), "workspace too small");
    const int z = 2 * g.d;
    std::vector<PottsTab> tabs(n_rep);
    for (int r = 0; r < n_rep; ++r) {
        CM_REQUIRE(std::isfinite(betas[r]), "beta must be finite");
        for (int d = 0; d < 8; ++d) tabs[r].T[d] = d <= z ? potts_threshold(std::exp(-betas[r] * d)) : 0;
        for (int d = 0; d < 16; ++d) tabs[r].w[d] = d <= z ? std::exp(betas[r] * d) : 0.0;
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                tabs[r].w2[a + 8 * b][0] = tabs[r].w[a];
                tabs[r].w2[a + 8 * b][1] = tabs[r].w[b];
            }
    }
    CM_HIP_CHECK(upload_async(workspace, tabs.data(), tabs.size() * sizeof(PottsTab),
                              (hipStream_t)stream));
    return CM_OK;
}

// The sweeps of cm_potts_run with the tables already in the workspace: kernel
// launches only (n_meas = 0 validates the arguments and launches nothing).
int potts_rows(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
               const uint32_t* keys, const uint32_t* stream_ids, uint64_t* sweeps,
               const int64_t* schedule, int32_t n_rows, int64_t n_meas, int32_t* e_run,
               int32_t* counts_run, int32_t* e_out, int32_t* acc_out, int32_t* counts_out,
               void* workspace, int64_t workspace_bytes, int32_t block, int32_t grid,
               const int64_t* n_meas_rep, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(q >= 2 && q <= POTTS_QMAX, "q must be in [2, 127], got %d", q);
    CM_REQUIRE(n_rep >= 1, "n_rep must be positive");
    CM_REQUIRE(n_meas >= 0, "n_meas must not be negative");
    CM_REQUIRE(n_rows >= 1 && n_rows <= 8, "schedule must have 1..8 rows");
    CM_REQUIRE(workspace_bytes >= cm_potts_workspace_bytes(n_rep), "workspace too small");
    CM_REQUIRE(g.N <= (int64_t)1 << 30, "lattice too large");
    Sched sc{};
    sc.n = n_rows;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r], count = schedule[2 * r + 1];
        CM_REQUIRE(code == CODE_METROPOLIS || code == CODE_HEATBATH,
                   "update code %lld has no checkerboard Potts kernel", (long long)code);
        CM_REQUIRE(count >= 1 && count <= (1 << 20), "repetition count out of range");
        sc.code[r] = (int32_t)code;
        sc.count[r] = (int32_t)count;
    }
    if (n_meas == 0) return CM_OK;
    const int z = 2 * g.d;
    hipStream_t st = (hipStream_t)stream;
    PottsTab* dtabs = (PottsTab*)workspace;
    if (block <= 0) {  // one lane per site quad of a colour, in whole waves, at most 256
        const int64_t quads = (g.half + 3) / 4;
        block = (int)std::min<int64_t>(256, std::max<int64_t>(64, (quads + 63) / 64 * 64));
    }
    CM_REQUIRE(block % 64 == 0 && block <= 256, "block must be a multiple of 64, <= 256");
    // grid < 0 asks for the wide form, grid > 0 for one workgroup per replica;
    // 0 picks by potts_prefer_wide.
    const bool wide = grid < 0 || (grid == 0 && potts_prefer_wide(g.N, n_rep));
    if (wide) {
        // wide form: the whole GPU on each half sweep of every replica
        CM_REQUIRE(n_rep <= 65535, "n_rep out of range");
        CM_REQUIRE((g.half + 1023) / 1024 * n_rep < (1ll << 32), "too many sites x replicas per launch");
        const int64_t total = (int64_t)n_rep * n_meas;
        const int zb = (int)std::min<int64_t>((total * (q + 2) + 255) / 256, 8192);
        hipLaunchKernelGGL(potts_zero, dim3(zb), dim3(256), 0, st, e_out, total);
        hipLaunchKernelGGL(potts_zero, dim3(zb), dim3(256), 0, st, acc_out, total);
        hipLaunchKernelGGL(potts_zero, dim3(zb), dim3(256), 0, st, counts_out, total * q);
        const int64_t nquads = (g.half + 3) / 4;
        // quads per lane: up to 4 while the launch keeps >= 2048 waves (8 per CU)
        const int qpl = (int)std::max<int64_t>(1, std::min<int64_t>(4, (int64_t)n_rep * nquads / (64 * 2048)));
        const dim3 wgrid((unsigned)((nquads + 256 * qpl - 1) / (256 * qpl)), n_rep);
        uint64_t sw = 0;
        int64_t per_meas = 0;
        for (int r = 0; r < sc.n; ++r) per_meas += sc.count[r];
        for (int64_t meas = 0; meas < n_meas; ++meas)
            for (int r = 0; r < sc.n; ++r)
                for (int rp = 0; rp < sc.count[r]; ++rp) {
                    for (int c = 0; c < 2; ++c) {
                        if (z == 2)
                            launch_wide<2>(wgrid, st, (int8_t*)spins, g, q, sc.code[r], c, sw, dtabs, keys, stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out, acc_out, counts_out, qpl);
                        else if (z == 4)
                            launch_wide<4>(wgrid, st, (int8_t*)spins, g, q, sc.code[r], c, sw, dtabs, keys, stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out, acc_out, counts_out, qpl);
                        else
                            launch_wide<6>(wgrid, st, (int8_t*)spins, g, q, sc.code[r], c, sw, dtabs, keys, stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out, acc_out, counts_out, qpl);
                    }
                    ++sw;
                }
        hipLaunchKernelGGL(potts_prefix, dim3((unsigned)(((int64_t)n_rep * (q + 1) + 255) / 256)),
                           dim3(256), 0, st, e_out,
                           counts_out, n_meas, e_run, counts_run, q, n_rep, n_meas_rep, sweeps,
                           (uint64_t)per_meas);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    }
    if (grid <= 0) grid = n_rep;
    grid = std::min(grid, n_rep);
    hipError_t err;
    int8_t* s = (int8_t*)spins;
    if (z == 2)
        err = launch_potts_q<2>(s, g, q, n_rep, dtabs, keys, stream_ids, sweeps, sc, n_meas, e_run,
                              counts_run, e_out, acc_out, counts_out, n_meas_rep, block, grid, st);
    else if (z == 4)
        err = launch_potts_q<4>(s, g, q, n_rep, dtabs, keys, stream_ids, sweeps, sc, n_meas, e_run,
                              counts_run, e_out, acc_out, counts_out, n_meas_rep, block, grid, st);
    else
        err = launch_potts_q<6>(s, g, q, n_rep, dtabs, keys, stream_ids, sweeps, sc, n_meas, e_run,
                              counts_run, e_out, acc_out, counts_out, n_meas_rep, block, grid, st);
    CM_HIP_CHECK(err);
    return CM_OK;
}

}  // namespace cm

extern "C" {

int64_t cm_potts_workspace_bytes(int32_t n_rep) { return (int64_t)n_rep * sizeof(PottsTab); }

int cm_potts_init(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
                  const uint32_t* keys, const uint32_t* stream_ids, int32_t start, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(q >= 2 && q <= POTTS_QMAX, "q must be in [2, 127], got %d", q);
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    CM_REQUIRE(start == 0 || start == 1, "start must be 0 (random) or 1 (ordered)");
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>((g.N / 4 + 255) / 256, 1), 1024);
    hipLaunchKernelGGL(potts_init_kernel, dim3((unsigned)blocks, n_rep), dim3(256), 0,
                       (hipStream_t)stream, (int8_t*)spins, g.N, n_rep, q, keys, stream_ids,
                       start);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_potts_measure(const void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                     int32_t q, int32_t* e_run, int32_t* counts_run, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(q >= 2 && q <= POTTS_QMAX, "q must be in [2, 127], got %d", q);
    hipStream_t st = (hipStream_t)stream;
    const int8_t* s = (const int8_t*)spins;
    int64_t chunk = 4096;
    while ((int64_t)n_rep * ((g.N + chunk - 1) / chunk) > (int64_t)1 << 22) chunk *= 2;
    const int64_t nchunk = (g.N + chunk - 1) / chunk;
    const dim3 grid((unsigned)(n_rep * nchunk));
    hipLaunchKernelGGL(potts_measure_zero, dim3((unsigned)(((int64_t)n_rep * q + 255) / 256)),
                       dim3(256), 0, st, e_run, counts_run, n_rep, q);
    if (g.d == 1)
        hipLaunchKernelGGL(potts_measure_kernel<2>, grid, dim3(256), 0, st, s, g, q, chunk, nchunk,
                           e_run, counts_run);
    else if (g.d == 2)
        hipLaunchKernelGGL(potts_measure_kernel<4>, grid, dim3(256), 0, st, s, g, q, chunk, nchunk,
                           e_run, counts_run);
    else
        hipLaunchKernelGGL(potts_measure_kernel<6>, grid, dim3(256), 0, st, s, g, q, chunk, nchunk,
                           e_run, counts_run);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_potts_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
                 const double* betas, const uint32_t* keys, const uint32_t* stream_ids,
                 uint64_t* sweeps, const int64_t* schedule, int32_t n_rows, int64_t n_meas,
                 int32_t* e_run, int32_t* counts_run, int32_t* e_out, int32_t* acc_out,
                 int32_t* counts_out, void* workspace, int64_t workspace_bytes, int32_t block,
                 int32_t grid, const int64_t* n_meas_rep, void* stream)
{
    if (int rc = potts_rows(spins, ndim, dims, n_rep, q, keys, stream_ids, sweeps, schedule,
                            n_rows, 0, e_run, counts_run, e_out, acc_out, counts_out, workspace,
                            workspace_bytes, block, grid, n_meas_rep, stream))
        return rc;  // arguments checked before anything is uploaded
    if (n_meas == 0) return CM_OK;
    if (betas != nullptr)  // null: cm_potts_prepare already put these chains' tables in the workspace
        if (int rc = potts_prepare(ndim, dims, n_rep, betas, workspace, workspace_bytes, stream))
            return rc;
    return potts_rows(spins, ndim, dims, n_rep, q, keys, stream_ids, sweeps, schedule, n_rows,
                      n_meas, e_run, counts_run, e_out, acc_out, counts_out, workspace,
                      workspace_bytes, block, grid, n_meas_rep, stream);
}

int cm_potts_prepare(int32_t ndim, const int32_t* dims, int32_t n_rep, const double* betas,
                     void* workspace, int64_t workspace_bytes, void* stream)
{
    CM_REQUIRE(betas != nullptr, "betas must not be null");
    return potts_prepare(ndim, dims, n_rep, betas, workspace, workspace_bytes, stream);
}

}  // extern "C"
