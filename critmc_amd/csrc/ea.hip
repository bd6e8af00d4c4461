// Edwards-Anderson spin glass with parallel tempering, batched over disorder
// samples of uneven size; +-J couplings (int8, exact int32 energies) or
// Gaussian couplings (FP64 couplings, fields and energies).
//
// Reference (critmc/kernels/glass.py, critmc/glass.py):
//   sweep     glass.py:41-54   h = sum_k J_ik s_k, dE = 2 s h; Metropolis accept iff
//                              dE <= 0 or u < exp(-beta dE); the heat bath takes
//                              P(+1) = 1/(1 + exp(-2 beta h)) (ising.py:59-65 with field h)
//   swaps     glass.py:129-137 k = 0..m-2 in order, x = (b_k - b_{k+1})(E_c1 - E_c2),
//                              accept iff x >= 0 or u < exp(x); only perm changes
//   labels    glass.py:139-152 hot end -> 1, cold end -> 2, one-way trips, label_hist
//   measure   glass.py:56-77, 153-163  q, q_l, |q(k_min)|^2, log-binned accumulators
//
// Layout: sample s owns int8 spins [2 sets][m configurations][N_s] at spin_off[s]
// and forward-bond couplings [d][N_s] (int8 +-1 or FP64) at j_off[s] (bond
// a*N + i joins i and its +e_a neighbour, critmc/lattice.py:136-138).  perm/kofc/labels/energies/
// sweep counters are [S][2][m] (configuration-indexed except perm, which maps
// temperature -> configuration).
//
// One PT step is three launches.  Per-configuration path (Gaussian
// couplings, or batches ea_use_ms rejects): ea_sweep (one workgroup per
// configuration, spins staged in LDS), ea_swap (one 64-lane workgroup per
// (sample, set), chain in LDS), ea_measure (one workgroup per (sample,
// temperature)).  Multi-spin path (+-J, m <= 32): the configurations are
// packed one bit each into words for the call (ea_ms_pack / ea_ms_unpack);
// ea_ms_sweep (one workgroup per (sample, set)), ea_swap, ea_measure_ms (one
// workgroup per sample).
//
// Uniforms (key = the sample's), per-configuration path: a configuration's
// site k of colour c uses word k&3 of block k>>2 at (sweep, c, TAG_EA,
// stream r*m + c).  Multi-spin path: see ea_ms_pack.  The swap
// between temperatures k and k+1 of set r uses word k&3 of block k>>2 at
// (step, 0, TAG_PT, stream r).  The swap test compares u = x 2^-32 with
// det_exp(x), a fixed sequence of IEEE operations (Cody-Waite reduction,
// degree-13 Taylor polynomial in fma form) evaluated identically on the host.
//
// Gaussian couplings: h = sum_k J_ik s_k is added in the reference's column
// order (+x, -x, +y, ...), one FP64 addition per neighbour; Metropolis accepts
// iff dE <= 0 or u < det_exp(-beta dE), the heat bath sets +1 iff
// u < 1 / (1 + det_exp(-2 beta h)), u = x 2^-32.  Spins, swaps and labels are
// therefore reproducible bit for bit on the host; the running FP64 energies
// are sums in the workgroup's reduction order, recomputed from the spins at
// the start of every bin after the first (critmc/kernels/glass.py:116-122).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "cb_geom.hpp"
#include "common.hpp"
#include "det_exp.hpp"
#include "ising_core.hpp"
#include "philox.hpp"

namespace cm {

// Multi-spin batches run as one fused launch per call (1) or three launches
// per PT step (0); both give the same chains.
int64_t g_ea_fused = 1;
int64_t g_ea_order = 0;

constexpr int EA_NOBS = 7;
constexpr int EA_TAB = 16;
constexpr int64_t EA_LDS_MAX = 32768;
constexpr int EA_LMAX = 256;

struct EaArgs {
    int8_t* spins;
    const int8_t* J;        // +-1 couplings, or
    const double* Jf;       // Gaussian couplings (one of the two is set)
    const int32_t* L;
    const int64_t* spin_off;
    const int64_t* j_off;
    const uint32_t* keys;   // [S][2]
    const double* betas;    // [S][m], hot -> cold
    const uint64_t* thr;    // [S][m][EA_TAB]
    uint64_t* sweeps;       // [S][2][m]
    int32_t* energies;      // [S][2][m] (+-J)
    double* energies_f;     // [S][2][m] (Gaussian)
    int32_t* perm;          // [S][2][m]
    int32_t* kofc;          // [S][2][m]
    int32_t* labels;        // [S][2][m]
    int32_t* trips;         // [S][2]
    int32_t* label_hist;    // [S][2][m][2]
    int32_t* swap_acc;      // [S][2][m-1]
    int32_t* swap_try;      // [S][2][m-1]
    double* acc;            // [S][n_bins][m][EA_NOBS]
    int32_t* e_series;      // [S][2][m][n_series] (temperature-indexed) or null
    double* e_series_f;     // the same for Gaussian couplings
    int32_t* q_series;      // [S][m][n_series] or null
    // multi-spin +-J path (m <= 32): one bit per configuration
    uint32_t* msw;          // [sum_s 2 N_s] words: sample s, set r, site i at ms_off[s] + r N_s + i
    uint8_t* jneg;          // [sum_s N_s]: bit j set where the coupling to neighbour j is -1
    const int64_t* ms_off;  // [S] word offsets (the jneg offset is ms_off[s] / 2)
    const int32_t* ms_order;  // [2 S] (sample, set) = (v >> 1, v & 1) per workgroup, largest first
    const int32_t* samples;   // [S] fused path: the sample of each workgroup
    int S, m, d, code, sweeps_per_step, n_bins;
    int64_t n_series;
};

__device__ __forceinline__ CbGeo ea_geo(int d, int L)
{
    CbGeo g;
    g.d = d;
    g.L0 = L;
    g.L1 = d >= 2 ? L : 1;
    g.L2 = d >= 3 ? L : 1;
    g.M = L / 2;
    g.N = (int64_t)L * g.L1 * g.L2;
    g.half = g.N / 2;
    cb_finish(g);
    return g;
}

// Coordinates of site i (x fastest) on an L^d lattice, 32-bit arithmetic.
template <int D>
__device__ __forceinline__ void ea_coords(uint32_t i, uint32_t L, const FastDiv& dl, int (&x)[3])
{
    uint32_t rest = i;
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
        const uint32_t q = fdiv(rest, dl);
        x[ax] = (int)(rest - q * L);
        rest = q;
    }
}

// LDS: every sample's configuration fits EA_LDS_MAX bytes and is staged in
// LDS for the sweep (so spin reads compile to LDS loads); otherwise the sweep
// works on global memory.
template <int Z, bool G, bool LDS>
__global__ __launch_bounds__(256) void ea_sweep(EaArgs a)
{
    extern __shared__ int8_t lds8[];
    __shared__ int red[16];
    __shared__ double redf[4];
    const int m = a.m;
    const int s = blockIdx.x / (2 * m);
    const int rc = blockIdx.x - s * 2 * m;  // r * m + c
    const CbGeo g = ea_geo(Z / 2, a.L[s]);
    const int64_t N = g.N;
    int8_t* gs = a.spins + a.spin_off[s] + (int64_t)rc * N;
    const int8_t* J = G ? nullptr : a.J + a.j_off[s];
    const double* Jf = G ? a.Jf + a.j_off[s] : nullptr;
    int8_t* S = LDS ? lds8 : gs;
    if (LDS)
        for (int64_t t = threadIdx.x; t < N; t += blockDim.x) lds8[t] = gs[t];
    const int64_t cfg = (int64_t)s * 2 * m + rc;
    const int k = a.kofc[cfg];
    const uint64_t* T = a.thr + ((int64_t)s * m + k) * EA_TAB;
    uint64_t sweep = a.sweeps[cfg];
    const uint32_t k0 = a.keys[2 * s], k1 = a.keys[2 * s + 1], sid = (uint32_t)rc;
    const double beta = a.betas[(int64_t)s * m + k];
    __syncthreads();
    int de = 0;
    double def = 0.0;
    const int64_t nquad = (g.half + 3) >> 2;
    for (int sp = 0; sp < a.sweeps_per_step; ++sp) {
        for (int c = 0; c < 2; ++c) {
            const uint32_t zc = ctr_z(sweep, (uint32_t)c, TAG_EA);
            for (int64_t qd = threadIdx.x; qd < nquad; qd += blockDim.x) {
                const u32x4 rn = philox4x32_10(u32x4{(uint32_t)qd, (uint32_t)sweep, zc, sid}, k0, k1);
                const uint32_t xs[4] = {rn.x, rn.y, rn.z, rn.w};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int64_t kk = 4 * qd + t;
                    if (kk >= g.half) break;
                    uint32_t nb[Z];
                    const uint32_t i = cb_site32<Z>(g, (uint32_t)kk, c, nb);
                    const uint32_t N32 = (uint32_t)N;
                    const int si = S[i];
                    if constexpr (G) {
                        double h = 0.0;
#pragma unroll
                        for (int ax = 0; ax < Z / 2; ++ax) {
                            h += Jf[ax * N32 + i] * (double)S[nb[2 * ax]];
                            h += Jf[ax * N32 + nb[2 * ax + 1]] * (double)S[nb[2 * ax + 1]];
                        }
                        const double u = (double)xs[t] * 0x1p-32;
                        if (a.code == CODE_METROPOLIS) {
                            const double dE = 2.0 * (double)si * h;
                            if (dE <= 0.0 || u < det_exp(-beta * dE)) {
                                S[i] = (int8_t)(-si);
                                def += dE;
                            }
                        } else {
                            const int nw = u < 1.0 / (1.0 + det_exp(-2.0 * beta * h)) ? 1 : -1;
                            if (nw != si) {
                                def += 2.0 * (double)si * h;
                                S[i] = (int8_t)nw;
                            }
                        }
                    } else {
                        int h = 0;
#pragma unroll
                        for (int ax = 0; ax < Z / 2; ++ax) {
                            h += J[ax * N32 + i] * S[nb[2 * ax]];
                            h += J[ax * N32 + nb[2 * ax + 1]] * S[nb[2 * ax + 1]];
                        }
                        if (a.code == CODE_METROPOLIS) {
                            const int aa = si * h;
                            if (aa <= 0 || (uint64_t)xs[t] < T[aa + Z]) {
                                S[i] = (int8_t)(-si);
                                de += 2 * aa;
                            }
                        } else {
                            const int nw = (uint64_t)xs[t] < T[h + Z] ? 1 : -1;
                            if (nw != si) {
                                de += 2 * si * h;
                                S[i] = (int8_t)nw;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
        ++sweep;
    }
    if constexpr (G) {
        def = wave_sum_f64(def);
        if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = def;
    } else {
        de = wave_sum(de);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = de;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)((blockDim.x + 63) >> 6);
        if constexpr (G) {
            double t = 0.0;
            for (int w = 0; w < nw; ++w) t += redf[w];
            a.energies_f[cfg] += t;
        } else {
            int t = 0;
            for (int w = 0; w < nw; ++w) t += red[w];
            a.energies[cfg] += t;
        }
        a.sweeps[cfg] = sweep;
    }
    if (LDS)
        for (int64_t t = threadIdx.x; t < N; t += blockDim.x) gs[t] = lds8[t];
}

// ---------------------------------------------------------------------------
// Multi-spin +-J path (m <= 32 temperatures).  Word i of set r of sample s
// holds bit c = spin of configuration c (1 = +1) at site i, for every
// configuration of the set, so one 32-bit operation updates a site of all of
// them.  Configurations never change lanes: a swap changes only which
// temperature a configuration sits at, and each lane reads its own
// temperature's thresholds.  Colour site k draws the set's 32 bit planes from
// words 0..3 of Philox blocks 8k .. 8k+7 at (the set's sweep counter, colour,
// TAG_EA, stream r); configuration c's uniform is x = sum_t bit_c(plane_t)
// 2^(31-t), compared plane by plane with its threshold as in the Ising
// kernels.

// Pack int8 configurations into words; record the negative couplings of each
// site's 2d neighbours (column order +x, -x, +y, -y, +z, -z).
template <int D>
__global__ void ea_ms_pack(EaArgs a)
{
    const int s = blockIdx.y;
    const int L = a.L[s];
    const uint32_t N = (uint32_t)L * (D >= 2 ? L : 1) * (D >= 3 ? L : 1);
    const int m = a.m;
    const int8_t* sp = a.spins + a.spin_off[s];
    uint32_t* w = a.msw + a.ms_off[s];
    uint8_t* jn = a.jneg + a.ms_off[s] / 2;
    const int8_t* J = a.J + a.j_off[s];
    const FastDiv dl = make_fastdiv((uint32_t)L);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        for (int r = 0; r < 2; ++r) {
            uint32_t v = 0;
            for (int c = 0; c < m; ++c) v |= (sp[((int64_t)r * m + c) * N + i] > 0 ? 1u : 0u) << c;
            w[(int64_t)r * N + i] = v;
        }
        int x[3];
        ea_coords<D>(i, (uint32_t)L, dl, x);
        uint32_t bits = 0;
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            const uint32_t stride = ax == 0 ? 1u : (ax == 1 ? (uint32_t)L : (uint32_t)L * L);
            const uint32_t back = x[ax] == 0 ? i + (uint32_t)(L - 1) * stride : i - stride;
            bits |= (J[ax * N + i] < 0 ? 1u : 0u) << (2 * ax);
            bits |= (J[ax * N + back] < 0 ? 1u : 0u) << (2 * ax + 1);
        }
        jn[i] = (uint8_t)bits;
    }
}

__global__ void ea_ms_unpack(EaArgs a)
{
    const int s = blockIdx.y;
    const int L = a.L[s];
    const uint32_t N = (uint32_t)L * (a.d >= 2 ? L : 1) * (a.d >= 3 ? L : 1);
    const int m = a.m;
    int8_t* sp = a.spins + a.spin_off[s];
    const uint32_t* w = a.msw + a.ms_off[s];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        for (int r = 0; r < 2; ++r) {
            const uint32_t v = w[(int64_t)r * N + i];
            for (int c = 0; c < m; ++c) sp[((int64_t)r * m + c) * N + i] = ((v >> c) & 1u) ? 1 : -1;
        }
}

// out = V[c] per lane, c = c0 + 2 c1 + 4 c2 the lane's bit-sliced count.
__device__ __forceinline__ uint32_t ms_select(const uint32_t (&V)[8], uint32_t c0, uint32_t c1,
                                              uint32_t c2)
{
    const uint32_t a0 = (V[1] & c0) | (V[0] & ~c0), a1 = (V[3] & c0) | (V[2] & ~c0);
    const uint32_t a2 = (V[5] & c0) | (V[4] & ~c0), a3 = (V[7] & c0) | (V[6] & ~c0);
    const uint32_t b0 = (a1 & c1) | (a0 & ~c1), b1 = (a3 & c1) | (a2 & ~c1);
    return (b1 & c2) | (b0 & ~c2);
}

constexpr int EA_MS_CNT = 12;  // bit-sliced counter planes for the bond count of one lane
// Philox blocks a word may draw before leaving its wave's lockstep; words
// still undecided go to an LDS queue drained densely after the half sweep
// (the comparison continues unchanged, so every decision is the same).
constexpr uint32_t EA_MS_BUDGET = 2;
constexpr int EA_MS_QCAP = 1024;

// Colour site kk: its index, own word and the bit-sliced count of neighbours
// with J_ij s_j = +1 (heat bath) or aligned with s_i (Metropolis).
template <int Z>
__device__ __forceinline__ void ms_classify(const CbGeo& g, const uint32_t* __restrict__ W,
                                            const uint8_t* __restrict__ JN, uint32_t kk, int col,
                                            int code, uint32_t& i, uint32_t& own, uint32_t& c0,
                                            uint32_t& c1, uint32_t& c2)
{
    uint32_t nb[Z];
    i = cb_site32<Z>(g, kk, col, nb);
    own = W[i];
    const uint32_t jn = JN[i];
    uint32_t in[Z];
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        const uint32_t up = W[nb[j]] ^ (0u - ((jn >> j) & 1u));  // J_ij s_j = +1
        in[j] = code == CODE_METROPOLIS ? ~(own ^ up) : up;
    }
    count_bits<Z>(in, c0, c1, c2);
}

// Bit planes 4 q_begin .. 4 q_end - 1 of the comparison of every pending lane.
__device__ __forceinline__ void ms_draw(const uint32_t (*__restrict__ P)[8], uint32_t c0, uint32_t c1,
                                        uint32_t c2, uint32_t& pending, uint32_t& below,
                                        uint32_t q_begin, uint32_t q_end, uint32_t kk,
                                        uint64_t sweep, uint32_t zc, int r, uint32_t k0, uint32_t k1)
{
    for (uint32_t q = q_begin; q < q_end && pending != 0u; ++q) {
        const u32x4 rn = philox4x32_10(u32x4{8u * kk + q, (uint32_t)sweep, zc, (uint32_t)r}, k0, k1);
        const uint32_t planes[4] = {rn.x, rn.y, rn.z, rn.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint4* pr = (const uint4*)P[4 * q + t];
            const uint4 lo = pr[0], hi = pr[1];
            const uint32_t V[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            const uint32_t tb = ms_select(V, c0, c1, c2);
            const uint32_t diff = pending & (planes[t] ^ tb);
            below |= diff & tb;
            pending &= ~diff;
        }
    }
}

// One workgroup per (sample, set): every sweep of the step for all of the set's
// configurations, then each configuration's energy recounted from its bonds.
// LDS: the set's words and neighbour-coupling bytes when they fit, and the
// threshold planes P[t][class] (bit c = bit 31-t of configuration c's
// threshold for that neighbour class) with the always-accept masks.
#ifndef CM_EA_MS_THREADS
#define CM_EA_MS_THREADS 256
#endif
constexpr int EA_MS_THREADS = CM_EA_MS_THREADS;  // lanes per (sample, set) workgroup

template <int Z, bool LDS>
__global__ __launch_bounds__(EA_MS_THREADS) void ea_ms_sweep(EaArgs a)
{
    extern __shared__ uint32_t lds32[];
    __shared__ __align__(16) uint32_t P[32][8];
    __shared__ uint32_t ALW[8];
    __shared__ int ucnt[32];
    __shared__ uint32_t queue[3 * EA_MS_QCAP];  // colour site, pending, below
    __shared__ int qcount[2];                   // alternating half sweeps
    const int m = a.m;
    const int v = a.ms_order[blockIdx.x];
    const int s = v >> 1, r = v & 1;
    const CbGeo g = ea_geo(Z / 2, a.L[s]);
    const uint32_t N = (uint32_t)g.N;
    uint32_t* gw = a.msw + a.ms_off[s] + (int64_t)r * N;
    const uint8_t* gj = a.jneg + a.ms_off[s] / 2;
    uint32_t* W = LDS ? lds32 : gw;
    const uint8_t* JN = LDS ? (const uint8_t*)(lds32 + N) : gj;
    if (LDS) {
        for (uint32_t t = threadIdx.x; t < N; t += blockDim.x) lds32[t] = gw[t];
        uint8_t* jl = (uint8_t*)(lds32 + N);
        for (uint32_t t = threadIdx.x; t < N; t += blockDim.x) jl[t] = gj[t];
    }
    const int64_t cfg0 = ((int64_t)s * 2 + r) * m;
    if (threadIdx.x < 32) ucnt[threadIdx.x] = 0;
    if (threadIdx.x < 2) qcount[threadIdx.x] = 0;
    if (threadIdx.x < 64) {
        // lane c < m: configuration c's thresholds T[2 cl] for neighbour class cl
        const int c = threadIdx.x;
        const uint64_t* T = c < m ? a.thr + ((int64_t)s * m + a.kofc[cfg0 + c]) * EA_TAB : nullptr;
        for (int cl = 0; cl < 8; ++cl) {
            const uint64_t tv = (c < m && cl <= Z) ? T[2 * cl] : 0;
            const uint64_t alw = __ballot(tv >= (1ull << 32));
            if (threadIdx.x == 0) ALW[cl] = (uint32_t)alw;
            for (int t = 0; t < 32; ++t) {
                const uint64_t b = __ballot(tv < (1ull << 32) && ((tv >> (31 - t)) & 1ull));
                if (threadIdx.x == 0) P[t][cl] = (uint32_t)b;
            }
        }
    }
    __syncthreads();
    const uint32_t lanes = m >= 32 ? ~0u : ((1u << m) - 1u);
    uint32_t alw[8];
#pragma unroll
    for (int cl = 0; cl < 8; ++cl) alw[cl] = ALW[cl];
    uint64_t sweep = a.sweeps[cfg0];
    const uint32_t k0 = a.keys[2 * s], k1 = a.keys[2 * s + 1];
    const uint32_t half = (uint32_t)g.half;
    uint32_t hs = 0;
    for (int sp = 0; sp < a.sweeps_per_step; ++sp) {
        for (int col = 0; col < 2; ++col, ++hs) {
            const uint32_t zc = ctr_z(sweep, (uint32_t)col, TAG_EA);
            int* qc = qcount + (hs & 1);
            for (uint32_t kk = threadIdx.x; kk < half; kk += blockDim.x) {
                uint32_t i, own, c0, c1, c2;
                ms_classify<Z>(g, W, JN, kk, col, a.code, i, own, c0, c1, c2);
                uint32_t below = ms_select(alw, c0, c1, c2) & lanes;
                uint32_t pending = lanes & ~below;
                ms_draw(P, c0, c1, c2, pending, below, 0, EA_MS_BUDGET, kk, sweep, zc, r, k0, k1);
                if (pending) {
                    const int slot = atomicAdd(qc, 1);
                    if (slot < EA_MS_QCAP) {
                        queue[3 * slot] = kk;
                        queue[3 * slot + 1] = pending;
                        queue[3 * slot + 2] = below;
                        continue;
                    }
                    ms_draw(P, c0, c1, c2, pending, below, EA_MS_BUDGET, 8, kk, sweep, zc, r, k0, k1);
                }
                W[i] = a.code == CODE_METROPOLIS ? own ^ below : below;
            }
            __syncthreads();
            const int nq = min(*qc, EA_MS_QCAP);
            if (threadIdx.x == 0) qcount[(hs + 1) & 1] = 0;
            for (int t = threadIdx.x; t < nq; t += blockDim.x) {
                const uint32_t kk = queue[3 * t];
                uint32_t i, own, c0, c1, c2;
                ms_classify<Z>(g, W, JN, kk, col, a.code, i, own, c0, c1, c2);
                uint32_t pending = queue[3 * t + 1], below = queue[3 * t + 2];
                ms_draw(P, c0, c1, c2, pending, below, EA_MS_BUDGET, 8, kk, sweep, zc, r, k0, k1);
                W[i] = a.code == CODE_METROPOLIS ? own ^ below : below;
            }
            __syncthreads();
        }
        ++sweep;
    }
    // energies: E_c = 2 U_c - d N with U_c the configuration's unsatisfied bonds,
    // counted per lane in bit-sliced counters and reduced with ballots
    uint32_t cnt[EA_MS_CNT];
#pragma unroll
    for (int p = 0; p < EA_MS_CNT; ++p) cnt[p] = 0;
    const uint32_t L32 = (uint32_t)g.L0;
    const FastDiv dl = make_fastdiv(L32);
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
        int x[3];
        ea_coords<Z / 2>(i, L32, dl, x);
        const uint32_t si = W[i], jn = JN[i];
#pragma unroll
        for (int ax = 0; ax < Z / 2; ++ax) {
            const uint32_t stride = ax == 0 ? 1u : (ax == 1 ? L32 : L32 * L32);
            const uint32_t j = (uint32_t)x[ax] + 1 == L32 ? i - (L32 - 1) * stride : i + stride;
            uint32_t carry = si ^ W[j] ^ (0u - ((jn >> (2 * ax)) & 1u));  // J s_i s_j = -1
            for (int p = 0; p < EA_MS_CNT && carry; ++p) {
                const uint32_t t = cnt[p] & carry;
                cnt[p] ^= carry;
                carry = t;
            }
        }
    }
    // planes that can be non-zero: a lane counts at most d ceil(N / blockDim) bonds
    const uint32_t most = (uint32_t)(Z / 2) * ((N + blockDim.x - 1) / blockDim.x);
    const int planes = min(EA_MS_CNT, 32 - __clz(most));
    for (int c = 0; c < m; ++c) {
        int u = 0;
        for (int p = 0; p < planes; ++p) u += __popcll(__ballot((cnt[p] >> c) & 1u)) << p;
        if ((threadIdx.x & 63) == 0 && u) atomicAdd(&ucnt[c], u);
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)m) {
        a.energies[cfg0 + threadIdx.x] = 2 * ucnt[threadIdx.x] - (Z / 2) * (int)N;
        a.sweeps[cfg0 + threadIdx.x] = sweep;
    }
    if (LDS)
        for (uint32_t t = threadIdx.x; t < N; t += blockDim.x) gw[t] = lds32[t];
}

// One 64-lane workgroup per (sample, set): the lanes stage the set's
// permutation, labels, energies, betas and swap uniforms in LDS, lane 0 runs
// the sequential swap chain k = 0..m-2 out of LDS, and the lanes write back.
constexpr int EA_MMAX = 256;

template <bool G>
__global__ __launch_bounds__(64) void ea_swap(EaArgs a, uint64_t step)
{
    __shared__ int32_t sp[EA_MMAX], sk[EA_MMAX], sl[EA_MMAX], sacc_f[EA_MMAX];
    __shared__ double sE[EA_MMAX], sb[EA_MMAX];
    __shared__ uint32_t sx[EA_MMAX];
    __shared__ int s_trip;
    const int t = blockIdx.x;  // s * 2 + r
    const int s = t >> 1, r = t & 1;
    const int m = a.m;
    const int64_t base = (int64_t)t * m;
    int32_t* perm = a.perm + base;
    int32_t* kofc = a.kofc + base;
    int32_t* lab = a.labels + base;
    const uint32_t k0 = a.keys[2 * s], k1 = a.keys[2 * s + 1];
    const uint32_t zc = ctr_z(step, 0, TAG_PT);
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        sp[k] = perm[k];
        sk[k] = kofc[k];
        sl[k] = lab[k];
        sE[k] = G ? a.energies_f[base + k] : (double)a.energies[base + k];
        sb[k] = a.betas[(int64_t)s * m + k];
        sacc_f[k] = 0;
    }
    for (int b = threadIdx.x; 4 * b < m - 1; b += blockDim.x) {
        const u32x4 rn = philox4x32_10(u32x4{(uint32_t)b, (uint32_t)step, zc, (uint32_t)r}, k0, k1);
        sx[4 * b] = rn.x;
        if (4 * b + 1 < EA_MMAX) sx[4 * b + 1] = rn.y;
        if (4 * b + 2 < EA_MMAX) sx[4 * b + 2] = rn.z;
        if (4 * b + 3 < EA_MMAX) sx[4 * b + 3] = rn.w;
    }
    // m <= 32: the swaps are sequential, but the configuration at slot k when
    // pair (k, k+1) is tried can only be the one that started at some slot
    // j <= k (carried up by the swaps j .. k-1).  Every decision (j, k) is
    // taken in parallel first, with the same operations as the sequential
    // pass, which then only follows the chain: the same swaps bit for bit.
    __shared__ uint32_t dmask[32];
    const bool par = m <= 32;
    if (par) {
        if (threadIdx.x < 32) dmask[threadIdx.x] = 0u;
        __syncthreads();
        const int npair = (m - 1) * m / 2;  // (j, k), 0 <= j <= k < m - 1
        for (int p = threadIdx.x; p < npair; p += blockDim.x) {
            int k = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
            while ((k + 1) * (k + 2) / 2 <= p) ++k;
            while (k * (k + 1) / 2 > p) --k;
            const int j = p - k * (k + 1) / 2;
            const int c1 = sp[j], c2 = sp[k + 1];
            const double x = (sb[k] - sb[k + 1]) * (sE[c1] - sE[c2]);
            if (x >= 0.0 || (double)sx[k] * 0x1p-32 < det_exp(x)) atomicOr(&dmask[k], 1u << j);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int j = 0;  // slot where the configuration now at slot k started
        for (int k = 0; k < m - 1; ++k) {
            const int c1 = sp[k], c2 = sp[k + 1];
            bool swap;
            if (par) {
                swap = ((dmask[k] >> j) & 1u) != 0u;
            } else {
                const double x = (sb[k] - sb[k + 1]) * (sE[c1] - sE[c2]);
                swap = x >= 0.0 || (double)sx[k] * 0x1p-32 < det_exp(x);
            }
            if (swap) {
                sp[k] = c2;
                sp[k + 1] = c1;
                sk[c2] = k;
                sk[c1] = k + 1;
                sacc_f[k] = 1;
            } else {
                j = k + 1;
            }
        }
        const int hot = sp[0], cold = sp[m - 1];
        int trips = 0;
        if (sl[hot] == 2) ++trips;
        sl[hot] = 1;
        if (sl[cold] == 1) ++trips;
        sl[cold] = 2;
        s_trip = trips;
    }
    __syncthreads();
    int32_t* sacc = a.swap_acc + (int64_t)t * (m - 1);
    int32_t* stry = a.swap_try + (int64_t)t * (m - 1);
    int32_t* hist = a.label_hist + (int64_t)t * m * 2;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        perm[k] = sp[k];
        kofc[k] = sk[k];
        lab[k] = sl[k];
        if (k < m - 1) {
            stry[k] += 1;
            sacc[k] += sacc_f[k];
        }
        const int l = sl[sp[k]];
        if (l == 1) hist[2 * k] += 1;
        else if (l == 2) hist[2 * k + 1] += 1;
    }
    if (threadIdx.x == 0 && s_trip) a.trips[t] += s_trip;
}

// LDS: the site products q_i = s_a[i] s_b[i] of the two sets' configurations
// at this temperature are staged in LDS (16-byte loads), and q_l uses
// s_a[j] s_b[j] = q_j, so the site loop reads only that array.
template <int Z, bool G, bool LDS>
__global__ __launch_bounds__(256) void ea_measure(EaArgs a, int64_t slot)
{
    __shared__ int ired[2 * 16];
    __shared__ double dred[6 * 16];
    __shared__ double ctab[EA_LMAX], stab[EA_LMAX];
    const int m = a.m;
    const int s = blockIdx.x / m, k = blockIdx.x - (blockIdx.x / m) * m;
    const int L = a.L[s];
    const CbGeo g = ea_geo(Z / 2, L);
    const int64_t N = g.N;
    const int ca = a.perm[((int64_t)s * 2 + 0) * m + k];
    const int cb = a.perm[((int64_t)s * 2 + 1) * m + k];
    const int8_t* sa = a.spins + a.spin_off[s] + ((int64_t)0 * m + ca) * N;
    const int8_t* sb = a.spins + a.spin_off[s] + ((int64_t)1 * m + cb) * N;
    // cos / sin of 2 pi x / L for every coordinate value (L <= EA_LMAX)
    for (int x = threadIdx.x; x < L; x += blockDim.x) {
        double sn, cs;
        sincospi(2.0 * x / L, &sn, &cs);
        ctab[x] = cs;
        stab[x] = sn;
    }
    extern __shared__ int8_t qlds[];
    if constexpr (LDS) {
        // N = L^d with L even and d >= 2 is a multiple of 4, and every
        // configuration starts at a multiple of N: aligned 4-byte chunks
        const uint32_t* va = (const uint32_t*)sa;
        const uint32_t* vb = (const uint32_t*)sb;
        uint32_t* vq = (uint32_t*)qlds;
        for (uint32_t t = threadIdx.x; t < (uint32_t)(N >> 2); t += blockDim.x) {
            // +-1 bytes: the product is +1 (0x01) where the bytes agree and -1
            // (0xFF) where they differ; x ^ y is 0x00 or 0xFE per byte
            vq[t] = 0x01010101u | (va[t] ^ vb[t]);
        }
    }
    __syncthreads();
    const FastDiv dl = make_fastdiv((uint32_t)L);
    int q = 0, ql = 0;
    double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
    const uint32_t L32 = (uint32_t)L;
    for (uint32_t i = threadIdx.x; i < (uint32_t)N; i += blockDim.x) {
        const int qi = LDS ? qlds[i] : sa[i] * sb[i];
        q += qi;
        int x[3];
        ea_coords<Z / 2>(i, L32, dl, x);
#pragma unroll
        for (int ax = 0; ax < Z / 2; ++ax) {
            const uint32_t stride = ax == 0 ? 1u : (ax == 1 ? L32 : L32 * L32);
            const uint32_t j = (uint32_t)x[ax] + 1 == L32 ? i - (L32 - 1) * stride : i + stride;
            ql += qi * (LDS ? qlds[j] : sa[j] * sb[j]);
            re[ax] += qi * ctab[x[ax]];
            im[ax] += qi * stab[x[ax]];
        }
    }
    q = wave_sum(q);
    ql = wave_sum(ql);
    for (int ax = 0; ax < 3; ++ax) {
        re[ax] = wave_sum_f64(re[ax]);
        im[ax] = wave_sum_f64(im[ax]);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ired[2 * w] = q;
        ired[2 * w + 1] = ql;
        for (int ax = 0; ax < 3; ++ax) {
            dred[6 * w + ax] = re[ax];
            dred[6 * w + 3 + ax] = im[ax];
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int nw = (blockDim.x + 63) >> 6;
    int Q = 0, QL = 0;
    double RE[3] = {0, 0, 0}, IM[3] = {0, 0, 0};
    for (int v = 0; v < nw; ++v) {
        Q += ired[2 * v];
        QL += ired[2 * v + 1];
        for (int ax = 0; ax < 3; ++ax) {
            RE[ax] += dred[6 * v + ax];
            IM[ax] += dred[6 * v + 3 + ax];
        }
    }
    const int d = Z / 2;
    double qk2 = 0.0;
    for (int ax = 0; ax < d; ++ax) qk2 += (RE[ax] * RE[ax] + IM[ax] * IM[ax]) / ((double)N * (double)N);
    qk2 /= d;
    const double Ea = G ? a.energies_f[((int64_t)s * 2 + 0) * m + ca]
                        : (double)a.energies[((int64_t)s * 2 + 0) * m + ca];
    const double Eb = G ? a.energies_f[((int64_t)s * 2 + 1) * m + cb]
                        : (double)a.energies[((int64_t)s * 2 + 1) * m + cb];
    const int bin = 63 - __clzll((long long)(slot + 1));  // steps of this call, from 0
    if (bin < a.n_bins) {
        double* o = a.acc + (((int64_t)s * a.n_bins + bin) * m + k) * EA_NOBS;
        const double qq = (double)Q / (double)N;
        o[0] += 0.5 * (Ea + Eb) / (double)N;
        o[1] += (double)QL / (double)(d * N);
        o[2] += fabs(qq);
        o[3] += qq * qq;
        o[4] += qq * qq * qq * qq;
        o[5] += qk2;
        o[6] += 1.0;
    }
    if (G && a.e_series_f) {
        a.e_series_f[(((int64_t)s * 2 + 0) * m + k) * a.n_series + slot] = Ea;
        a.e_series_f[(((int64_t)s * 2 + 1) * m + k) * a.n_series + slot] = Eb;
    } else if (!G && a.e_series) {
        a.e_series[(((int64_t)s * 2 + 0) * m + k) * a.n_series + slot] = (int32_t)Ea;
        a.e_series[(((int64_t)s * 2 + 1) * m + k) * a.n_series + slot] = (int32_t)Eb;
    }
    if (a.q_series) a.q_series[((int64_t)s * m + k) * a.n_series + slot] = Q;
}

// 32 x 32 bit transpose in registers: bit j of v[i] becomes bit i of v[j].
__device__ __forceinline__ void transpose32(uint32_t (&v)[32])
{
    uint32_t msk = 0x0000FFFFu;
#pragma unroll
    for (int j = 16; j != 0; j >>= 1, msk ^= msk << j) {
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            if (k & j) continue;
            const uint32_t t = ((v[k] >> j) ^ v[k + j]) & msk;
            v[k + j] ^= t;
            v[k] ^= t << j;
        }
    }
}

// Transpose of 16 words as two 16 x 16 bit matrices side by side: bit j of
// v[i] (j < 16) becomes bit i of v[j], bit 16 + j becomes bit 16 + i of v[j].
// With v[i] the words of a row of at most 16 sites, v[c] & 0xFFFF is then
// configuration c's row pattern and v[c] >> 16 that of configuration c + 16
// (the 32 x 32 transpose without its first stage, whose halves are zero).
__device__ __forceinline__ void transpose16x2(uint32_t (&v)[16])
{
    uint32_t msk = 0x00FF00FFu;
#pragma unroll
    for (int j = 8; j != 0; j >>= 1, msk ^= msk << j) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k & j) continue;
            const uint32_t t = ((v[k] >> j) ^ v[k + j]) & msk;
            v[k + j] ^= t;
            v[k] ^= t << j;
        }
    }
}

constexpr int EA_MS_ROWS = 256;  // lattice rows (L^(d-1)) a measurement workgroup holds

// Multi-spin measurement, one workgroup per sample, all temperatures.
// Phase 1 (a lane per lattice row x1, x2): the row's words of both sets are
// transposed, so configuration c's spins along the row form one L-bit
// pattern; temperature k's overlap pattern is then qp = ~(A_ca ^ B_cb) (bit x
// set where q_i = +1), stored in LDS QP[row][k].
// Phase 2 (eight lanes per temperature, each a band of rows): q = 2 popc - L
// per row; link products from popcounts of qp against its rotation (x) and
// the patterns of the next rows (y, z); the x Fourier sum 2 sum_{bit x set}
// e^{i 2 pi x / L} from nibble tables (sum_x e^{i 2 pi x / L} = 0); the row
// sums go to per-plane sums M1[k][x1], M2[k][x2] for the y and z terms.
template <int Z, bool ROW16>
__global__ __launch_bounds__(256) void ea_measure_ms(EaArgs a, int64_t slot)
{
    constexpr int D = Z / 2;
    __shared__ uint32_t QP[EA_MS_ROWS][33];  // padded: lanes on different rows hit different banks
    __shared__ double ctab[EA_LMAX], stab[EA_LMAX];
    __shared__ double CT[8][16], ST[8][16];
    __shared__ int M1[32][32], M2[32][32];
    __shared__ int pca[32], pcb[32], kca[32], kcb[32];
    const int m = a.m;
    const int s = a.ms_order[2 * blockIdx.x] >> 1;  // largest samples first
    const int L = a.L[s];
    const uint32_t L32 = (uint32_t)L;
    const uint32_t N = L32 * (D >= 2 ? L32 : 1u) * (D >= 3 ? L32 : 1u);
    const uint32_t rows = N / L32;
    const uint32_t rowmask = L32 >= 32 ? ~0u : ((1u << L32) - 1u);
    const uint32_t* wa = a.msw + a.ms_off[s];
    const uint32_t* wb = wa + N;
    for (int x = threadIdx.x; x < L; x += blockDim.x) {
        double sn, cs;
        sincospi(2.0 * x / L, &sn, &cs);
        ctab[x] = cs;
        stab[x] = sn;
    }
    for (int t = threadIdx.x; t < 32 * 32; t += blockDim.x) {
        M1[t >> 5][t & 31] = 0;
        M2[t >> 5][t & 31] = 0;
    }
    if (threadIdx.x < 32) {
        const int k = threadIdx.x;
        pca[k] = k < m ? a.perm[((int64_t)s * 2 + 0) * m + k] : 0;
        pcb[k] = k < m ? a.perm[((int64_t)s * 2 + 1) * m + k] : 0;
        kca[k] = k < m ? a.kofc[((int64_t)s * 2 + 0) * m + k] : 0;
        kcb[k] = k < m ? a.kofc[((int64_t)s * 2 + 1) * m + k] : 0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 8 * 16; t += blockDim.x) {
        const int j = t >> 4, v = t & 15;
        double c = 0.0, sn = 0.0;
        for (int b = 0; b < 4; ++b) {
            const int x = 4 * j + b;
            if (((v >> b) & 1) && x < L) {
                c += ctab[x];
                sn += stab[x];
            }
        }
        CT[j][v] = c;
        ST[j][v] = sn;
    }
    // phase 1
    for (uint32_t row = threadIdx.x; row < rows && ROW16; row += blockDim.x) {
        uint32_t v[16];
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = (uint32_t)x < L32 ? wa[row * L32 + x] : 0u;
        transpose16x2(v);
#pragma unroll
        for (int c = 0; c < 32; ++c)
            if (c < m) QP[row][kca[c]] = c < 16 ? v[c] & 0xFFFFu : v[c - 16] >> 16;
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = (uint32_t)x < L32 ? wb[row * L32 + x] : 0u;
        transpose16x2(v);
#pragma unroll
        for (int c = 0; c < 32; ++c)
            if (c < m) {
                const uint32_t pb = c < 16 ? v[c] & 0xFFFFu : v[c - 16] >> 16;
                QP[row][kcb[c]] = ~(QP[row][kcb[c]] ^ pb) & rowmask;
            }
    }
    for (uint32_t row = threadIdx.x; row < rows && !ROW16; row += blockDim.x) {
        uint32_t v[32];
        // set 0: configuration c's row pattern is v[c]; place it at its temperature
#pragma unroll
        for (int x = 0; x < 32; ++x) v[x] = (uint32_t)x < L32 ? wa[row * L32 + x] : 0u;
        transpose32(v);
#pragma unroll
        for (int c = 0; c < 32; ++c)
            if (c < m) QP[row][kca[c]] = v[c];
        // set 1: the overlap pattern with the set-0 configuration at the same temperature
#pragma unroll
        for (int x = 0; x < 32; ++x) v[x] = (uint32_t)x < L32 ? wb[row * L32 + x] : 0u;
        transpose32(v);
#pragma unroll
        for (int c = 0; c < 32; ++c)
            if (c < m) QP[row][kcb[c]] = ~(QP[row][kcb[c]] ^ v[c]) & rowmask;
    }
    __syncthreads();
    // phase 2: lane = 8 k + j
    const int k = threadIdx.x >> 3, j = threadIdx.x & 7;
    int q = 0, ql = 0;
    double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;  // two chains: even and odd nibbles
    if (k < m) {
        const int nib = (L + 3) >> 2;
        // row = x1 + L x2, carried forward by 8 rows per step (no division per row)
        uint32_t x1 = 0, x2 = 0;
        if (D >= 2) {
            x1 = (uint32_t)j;
            while (x1 >= L32) {
                x1 -= L32;
                ++x2;
            }
        }
        for (uint32_t row = j; row < rows; row += 8) {
            const uint32_t qp = QP[row][k];
            const int rs = 2 * __popc(qp) - L;
            q += rs;
            const uint32_t rot = ((qp >> 1) | (qp << (L32 - 1))) & rowmask;  // bit x: q at x + 1
            ql += 2 * __popc(~(qp ^ rot) & rowmask) - L;
            if (D >= 2) {
                const uint32_t ry = (x1 + 1 == L32 ? 0u : x1 + 1) + L32 * x2;
                ql += 2 * __popc(~(qp ^ QP[ry][k]) & rowmask) - L;
                atomicAdd(&M1[k][x1], rs);
            }
            if (D >= 3) {
                const uint32_t rz = x1 + L32 * (x2 + 1 == L32 ? 0u : x2 + 1);
                ql += 2 * __popc(~(qp ^ QP[rz][k]) & rowmask) - L;
                atomicAdd(&M2[k][x2], rs);
            }
            for (int t = 0; t < nib; t += 2) {
                const uint32_t v = (qp >> (4 * t)) & 15u;
                re0 += CT[t][v];
                im0 += ST[t][v];
                if (t + 1 < nib) {
                    const uint32_t w = (qp >> (4 * t + 4)) & 15u;
                    re1 += CT[t + 1][w];
                    im1 += ST[t + 1][w];
                }
            }
            if (D >= 2) {
                x1 += 8;
                while (x1 >= L32) {
                    x1 -= L32;
                    ++x2;
                }
            }
        }
    }
    re0 += re1;
    im0 += im1;
    // reduce the eight lanes of each temperature (fixed order)
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        q += __shfl_xor(q, off, 64);
        ql += __shfl_xor(ql, off, 64);
        re0 += __shfl_xor(re0, off, 64);
        im0 += __shfl_xor(im0, off, 64);
    }
    __syncthreads();
    if (j != 0 || k >= m) return;
    double RE[3] = {2.0 * re0, 0, 0}, IM[3] = {2.0 * im0, 0, 0};
    for (int v = 0; v < L; ++v) {
        if (D >= 2) {
            RE[1] += M1[k][v] * ctab[v];
            IM[1] += M1[k][v] * stab[v];
        }
        if (D >= 3) {
            RE[2] += M2[k][v] * ctab[v];
            IM[2] += M2[k][v] * stab[v];
        }
    }
    double qk2 = 0.0;
    for (int ax = 0; ax < D; ++ax) qk2 += (RE[ax] * RE[ax] + IM[ax] * IM[ax]) / ((double)N * (double)N);
    qk2 /= D;
    const int ca = pca[k], cb = pcb[k];
    const double Ea = (double)a.energies[((int64_t)s * 2 + 0) * m + ca];
    const double Eb = (double)a.energies[((int64_t)s * 2 + 1) * m + cb];
    const int bin = 63 - __clzll((long long)(slot + 1));  // steps of this call, from 0
    if (bin < a.n_bins) {
        double* o = a.acc + (((int64_t)s * a.n_bins + bin) * m + k) * EA_NOBS;
        const double qq = (double)q / (double)N;
        o[0] += 0.5 * (Ea + Eb) / (double)N;
        o[1] += (double)ql / (double)(D * N);
        o[2] += fabs(qq);
        o[3] += qq * qq;
        o[4] += qq * qq * qq * qq;
        o[5] += qk2;
        o[6] += 1.0;
    }
    if (a.e_series) {
        a.e_series[(((int64_t)s * 2 + 0) * m + k) * a.n_series + slot] = (int32_t)Ea;
        a.e_series[(((int64_t)s * 2 + 1) * m + k) * a.n_series + slot] = (int32_t)Eb;
    }
    if (a.q_series) a.q_series[((int64_t)s * m + k) * a.n_series + slot] = q;
}

// ---------------------------------------------------------------------------
// Fused multi-spin tempering: one workgroup per sample runs every PT step of
// the call in one launch.  Lanes 0-255 sweep set 0 and lanes 256-511 set 1;
// both sets' words and the neighbour-coupling bytes stay in LDS for the whole
// call, and the energy recount, the swap chains, the threshold planes and the
// measurement run between barriers of the same workgroup, so samples of
// different sizes never wait for each other.  Every operation that decides a
// spin, a swap or an integer observable is the one of the three-launch step
// above, in the same order, so chains and series are the same bit for bit;
// only the FP64 Fourier sums of a measurement are added in another order
// (sixteen lanes per temperature).
//
// Dynamic LDS: words [2][N], coupling bytes [N] (padded to 16 B), then one
// region used by the sweep queues [2][3 QCAP] during sweeps and by the
// measurement (overlap patterns QP[32][rowsP], row sums M1[32][32] and
// M2[32][16]) afterwards.
constexpr int EAF_THREADS = 512;
constexpr int EAF_SET = 256;  // lanes per set
constexpr int64_t EAF_UNION = 2 * 3 * EA_MS_QCAP * 4;
constexpr int EAF_CNT = 8;  // counter planes of the energy recount: a lane counts at most 3 * 16 bonds

// Overlap patterns of one configuration are EAF_QP16 (16-bit patterns, rows
// <= 256) or EAF_QP32 (32-bit patterns, rows <= 32) entries apart: fixed
// strides, so pattern addresses are immediate offsets, and sixteen lanes of
// each of four temperatures read 16-bit patterns from disjoint LDS banks
// (272 / 2 = 136 = 8 mod 32 banks per configuration).
constexpr uint32_t EAF_QP16 = 272;
constexpr uint32_t EAF_QP32 = 80;

__host__ __device__ inline int64_t eaf_union_need(bool row16)
{
    return 32 * (row16 ? 2 * (int64_t)EAF_QP16 : 4 * (int64_t)EAF_QP32) + 4 * (32 * 32 + 32 * 16);
}

// P[r][t][cl] bit c = bit 31-t of the threshold of configuration c's
// temperature for neighbour class cl (0 where that class always accepts),
// ALW[r][cl] bit c = the class always accepts; lane = (r, t, cl).
__device__ __forceinline__ void eaf_planes(int tid, int m, const int (*kofc)[32], const uint32_t (*TT)[8],
                                           const uint32_t* TA, uint32_t (*P)[32][8], uint32_t (*ALW)[8])
{
    const int r = tid >> 8, t = (tid >> 3) & 31, cl = tid & 7;
    uint32_t p = 0, al = 0;
    for (int c = 0; c < m; ++c) {
        const int k = kofc[r][c];
        p |= ((TT[k][cl] >> (31 - t)) & 1u) << c;
        al |= ((TA[k] >> cl) & 1u) << c;
    }
    P[r][t][cl] = p;
    if (t == 0) ALW[r][cl] = al;
}

// Row `row` of a set: word x of the row in v[x] (x < L, zero above).
template <int NW>
__device__ __forceinline__ void eaf_row_words(const uint32_t* __restrict__ W, uint32_t row, uint32_t L,
                                              uint32_t (&v)[NW])
{
    const uint32_t* src = W + row * L;
    if ((L & 3u) == 0u) {
#pragma unroll
        for (int x = 0; x < NW; x += 4) {
            uint4 u = (uint32_t)x < L ? *(const uint4*)(src + x) : uint4{0u, 0u, 0u, 0u};
            v[x] = u.x;
            v[x + 1] = u.y;
            v[x + 2] = u.z;
            v[x + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int x = 0; x < NW; x += 2) {
            uint2 u = (uint32_t)x < L ? *(const uint2*)(src + x) : uint2{0u, 0u};
            v[x] = u.x;
            v[x + 1] = u.y;
        }
    }
}

// At most 128 VGPRs, so two workgroups (16 waves) share a CU; what the
// register allocator then spills lies outside the sweep loops.
template <int Z, bool ROW16>
__global__ __launch_bounds__(EAF_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
void ea_ms_fused(EaArgs a, uint64_t step0, int64_t n_steps)
{
    constexpr int D = Z / 2;
    using QT = typename std::conditional<ROW16, uint16_t, uint32_t>::type;
    extern __shared__ __align__(16) uint32_t dyn[];
    __shared__ __align__(16) uint32_t P[2][32][8];
    __shared__ uint32_t ALW[2][8];
    __shared__ uint32_t TT[32][8];  // threshold of temperature k, class cl (0 if it always accepts)
    __shared__ uint32_t TA[32];     // bit cl: temperature k always accepts class cl
    __shared__ int ucnt[2][32];
    __shared__ int qcount[2][2];
    __shared__ int sperm[2][32], skofc[2][32], slab[2][32], sE[2][32];
    __shared__ uint32_t sx[2][32];
    __shared__ __align__(16) uint32_t dmask[2][32];
    __shared__ __align__(16) uint32_t sigma[8];
    __shared__ double sbeta[32];
    __shared__ double ctab[32], stab[32];
    __shared__ double sacc[32][EA_NOBS];
    __shared__ int shist[2][32][2], sswap[2][32][2];  // label histogram; swap accepts and tries
    __shared__ int strips[2];

    const int m = a.m;
    const int s = a.samples[blockIdx.x];
    const int L = a.L[s];
    const CbGeo g = ea_geo(D, L);
    const uint32_t N = (uint32_t)g.N;
    const uint32_t L32 = (uint32_t)L;
    const uint32_t rows = N / L32;  // <= 256 (<= 32 for 32-bit patterns): one row per lane
    constexpr uint32_t rowsP = ROW16 ? EAF_QP16 : EAF_QP32;
    const uint32_t rowmask = L32 >= 32 ? ~0u : ((1u << L32) - 1u);
    const int tid = threadIdx.x, r = tid >> 8, tl = tid & (EAF_SET - 1);
    uint32_t* const W0 = dyn;
    uint32_t* const W1 = dyn + N;
    uint32_t* const W = r ? W1 : W0;
    uint8_t* const JN = (uint8_t*)(dyn + 2 * N);
    uint32_t* const U = dyn + 2 * N + (N + 15) / 16 * 4;
    uint32_t* const queue = U + r * 3 * EA_MS_QCAP;
    QT* const QP = (QT*)U;
    int* const M1 = (int*)(U + 32 * rowsP * (uint32_t)sizeof(QT) / 4);
    int* const M2 = M1 + 32 * 32;

    uint32_t* const gw = a.msw + a.ms_off[s];
    const uint8_t* const gj = a.jneg + a.ms_off[s] / 2;
    for (uint32_t t = tid; t < 2 * N; t += EAF_THREADS) dyn[t] = gw[t];
    for (uint32_t t = tid; t < N; t += EAF_THREADS) JN[t] = gj[t];
    const int64_t cfg0 = ((int64_t)s * 2 + r) * m;
    if (tl < m) {
        sperm[r][tl] = a.perm[cfg0 + tl];
        skofc[r][tl] = a.kofc[cfg0 + tl];
        slab[r][tl] = a.labels[cfg0 + tl];
        sE[r][tl] = a.energies[cfg0 + tl];
        shist[r][tl][0] = shist[r][tl][1] = 0;
        sswap[r][tl][0] = sswap[r][tl][1] = 0;
    }
    if (tl < 32) ucnt[r][tl] = 0;
    if (tl < 2) qcount[r][tl] = 0;
    if (tid < 8) sigma[tid] = 0u;
    if (tl == 0) strips[r] = 0;
    if (tid < m) {
        sbeta[tid] = a.betas[(int64_t)s * m + tid];
        const uint64_t* T = a.thr + ((int64_t)s * m + tid) * EA_TAB;
        uint32_t al = 0;
        for (int cl = 0; cl < 8; ++cl) {
            const uint64_t tv = cl <= Z ? T[2 * cl] : 0;
            TT[tid][cl] = tv < (1ull << 32) ? (uint32_t)tv : 0u;
            al |= (tv >= (1ull << 32) ? 1u : 0u) << cl;
        }
        TA[tid] = al;
    }
    if (tid < L) {
        double sn, cs;
        sincospi(2.0 * tid / L, &sn, &cs);
        ctab[tid] = cs;
        stab[tid] = sn;
    }
    const int64_t acc0 = (int64_t)s * a.n_bins * m * EA_NOBS;
    for (int e = tid; e < m * EA_NOBS; e += EAF_THREADS) sacc[e / EA_NOBS][e % EA_NOBS] = a.acc[acc0 + e];
    uint64_t sweep = a.sweeps[cfg0];
    const uint32_t k0 = a.keys[2 * s], k1 = a.keys[2 * s + 1];
    __syncthreads();
    eaf_planes(tid, m, skofc, TT, TA, P, ALW);
    __syncthreads();

    const uint32_t lanes = m >= 32 ? ~0u : ((1u << m) - 1u);
    const uint32_t half = (uint32_t)g.half;
    const int npair = (m - 1) * m / 2;
    const FastDiv dl = make_fastdiv(L32);
    const int mk = tid >> 4, mj = tid & 15;  // measurement lane: temperature, row band
    int cur_bin = 0;
    uint32_t hs = 0;
    for (int64_t t = 0; t < n_steps; ++t) {
        const uint64_t step = step0 + (uint64_t)t;
        // lane-derived values are recomputed every step from an opaque copy of
        // the lane index, so the compiler keeps them out of registers across steps
        int tid_ = tid;
        asm volatile("" : "+v"(tid_));
        const int tid = tid_, r = tid >> 8, tl = tid & (EAF_SET - 1), mk = tid >> 4, mj = tid & 15;
        uint32_t* const W = r ? W1 : W0;
        uint32_t* const queue = U + r * 3 * EA_MS_QCAP;
        // sweeps of both sets
        uint32_t alw[8];
#pragma unroll
        for (int cl = 0; cl < 8; ++cl) alw[cl] = ALW[r][cl];
        for (int sp = 0; sp < a.sweeps_per_step; ++sp) {
            for (int col = 0; col < 2; ++col, ++hs) {
                const uint32_t zc = ctr_z(sweep, (uint32_t)col, TAG_EA);
                int* qc = &qcount[r][hs & 1];
                for (uint32_t kk = tl; kk < half; kk += EAF_SET) {
                    uint32_t i, own, c0, c1, c2;
                    ms_classify<Z>(g, W, JN, kk, col, a.code, i, own, c0, c1, c2);
                    uint32_t below = ms_select(alw, c0, c1, c2) & lanes;
                    uint32_t pending = lanes & ~below;
                    ms_draw(P[r], c0, c1, c2, pending, below, 0, EA_MS_BUDGET, kk, sweep, zc, r, k0, k1);
                    if (pending) {
                        const int slot = atomicAdd(qc, 1);
                        if (slot < EA_MS_QCAP) {
                            queue[3 * slot] = kk;
                            queue[3 * slot + 1] = pending;
                            queue[3 * slot + 2] = below;
                            continue;
                        }
                        ms_draw(P[r], c0, c1, c2, pending, below, EA_MS_BUDGET, 8, kk, sweep, zc, r, k0, k1);
                    }
                    W[i] = a.code == CODE_METROPOLIS ? own ^ below : below;
                }
                __syncthreads();
                const int nq = min(*qc, EA_MS_QCAP);
                if (tl == 0) qcount[r][(hs + 1) & 1] = 0;
                for (int e = tl; e < nq; e += EAF_SET) {
                    const uint32_t kk = queue[3 * e];
                    uint32_t i, own, c0, c1, c2;
                    ms_classify<Z>(g, W, JN, kk, col, a.code, i, own, c0, c1, c2);
                    uint32_t pending = queue[3 * e + 1], below = queue[3 * e + 2];
                    ms_draw(P[r], c0, c1, c2, pending, below, EA_MS_BUDGET, 8, kk, sweep, zc, r, k0, k1);
                    W[i] = a.code == CODE_METROPOLIS ? own ^ below : below;
                }
                __syncthreads();
            }
            ++sweep;
        }
        // energies recounted from the bonds: per lane a bit-sliced binary counter
        // of unsatisfied bonds for every configuration, fed by carry-save adders
        // (the D bonds of a site first, then the running planes), then ballots
        {
            uint32_t cnt[EAF_CNT];
#pragma unroll
            for (int p = 0; p < EAF_CNT; ++p) cnt[p] = 0;
            for (uint32_t i = tl; i < N; i += EAF_SET) {
                int x[3];
                ea_coords<D>(i, L32, dl, x);
                const uint32_t si = W[i], jn = JN[i];
                uint32_t b[D];
#pragma unroll
                for (int ax = 0; ax < D; ++ax) {
                    const uint32_t stride = ax == 0 ? 1u : (ax == 1 ? L32 : L32 * L32);
                    const uint32_t j = (uint32_t)x[ax] + 1 == L32 ? i - (L32 - 1) * stride : i + stride;
                    b[ax] = si ^ W[j] ^ (0u - ((jn >> (2 * ax)) & 1u));  // J s_i s_j = -1
                }
                uint32_t s1, c2;  // the site's count: s1 + 2 c2
                if constexpr (D == 3) {
                    const uint32_t x01 = b[0] ^ b[1];
                    s1 = x01 ^ b[2];
                    c2 = (b[0] & b[1]) | (x01 & b[2]);
                } else {
                    s1 = b[0] ^ b[1];
                    c2 = b[0] & b[1];
                }
                const uint32_t k1 = cnt[0] & s1;  // weight 2
                cnt[0] ^= s1;
                const uint32_t t2 = cnt[1] ^ c2;
                uint32_t carry = (cnt[1] & c2) | (t2 & k1);  // weight 4
                cnt[1] = t2 ^ k1;
#pragma unroll
                for (int p = 2; p < EAF_CNT; ++p) {
                    const uint32_t tt = cnt[p] & carry;
                    cnt[p] ^= carry;
                    carry = tt;
                }
            }
            const uint32_t most = (uint32_t)D * ((N + EAF_SET - 1) / EAF_SET);
            const int planes = 32 - __clz(most);
            for (int c = 0; c < m; ++c) {
                int u = 0;
#pragma unroll
                for (int p = 0; p < EAF_CNT; ++p)
                    if (p < planes) u += __popcll(__ballot((cnt[p] >> c) & 1u)) << p;
                if ((tid & 63) == 0 && u) atomicAdd(&ucnt[r][c], u);
            }
        }
        if (tl < 32) dmask[r][tl] = 0u;
        if (4 * tl < m - 1) {
            const uint32_t zpt = ctr_z(step, 0, TAG_PT);
            const u32x4 rn = philox4x32_10(u32x4{(uint32_t)tl, (uint32_t)step, zpt, (uint32_t)r}, k0, k1);
            sx[r][4 * tl] = rn.x;
            sx[r][4 * tl + 1] = rn.y;
            sx[r][4 * tl + 2] = rn.z;
            sx[r][4 * tl + 3] = rn.w;
        }
        __syncthreads();
        if (tl < m) {
            sE[r][tl] = 2 * ucnt[r][tl] - D * (int)N;
            ucnt[r][tl] = 0;
        }
        __syncthreads();
        // every swap decision (j, k) in parallel, then the chain (as in ea_swap)
        for (int p = tl; p < npair; p += EAF_SET) {
            int k = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
            while ((k + 1) * (k + 2) / 2 <= p) ++k;
            while (k * (k + 1) / 2 > p) --k;
            const int j = p - k * (k + 1) / 2;
            const int c1 = sperm[r][j], c2 = sperm[r][k + 1];
            const double x = (sbeta[k] - sbeta[k + 1]) * ((double)sE[r][c1] - (double)sE[r][c2]);
            if (x >= 0.0 || (double)sx[r][k] * 0x1p-32 < det_exp(x)) atomicOr(&dmask[r][k], 1u << j);
        }
        __syncthreads();
        // The chain (as in ea_swap) in registers: flag k = swap (k, k+1) taken, the
        // decision of the configuration that started at slot j, carried to k by
        // the swaps j .. k-1.  A run of taken swaps a .. b moves the configuration
        // of slot a to b + 1 and those of a+1 .. b+1 down by one, so each slot's
        // new configuration follows from the flags.
        uint32_t flags = 0;
        {
            uint32_t dm[32];
#pragma unroll
            for (int k = 0; k < 32; k += 4) *(uint4*)(dm + k) = *(const uint4*)&dmask[r][k];
            int j = 0;
#pragma unroll
            for (int k = 0; k < 31; ++k) {
                if (k < m - 1) {
                    const uint32_t f = (dm[k] >> j) & 1u;
                    flags |= f << k;
                    j = f ? j : k + 1;
                }
            }
        }
        auto slot_config = [&](int k) {
            if ((flags >> k) & 1u) return sperm[r][k + 1];
            if (k >= 1 && ((flags >> (k - 1)) & 1u)) {
                const uint32_t zeros = ~flags & ((1u << (k - 1)) - 1u);
                return sperm[r][zeros ? 32 - __clz(zeros) : 0];
            }
            return sperm[r][k];
        };
        const int newc = tl < m ? slot_config(tl) : 0;
        if (tl == 0) {
            const int hot = newc, cold = slot_config(m - 1);
            int trips = 0;
            if (slab[r][hot] == 2) ++trips;
            slab[r][hot] = 1;
            if (slab[r][cold] == 1) ++trips;
            slab[r][cold] = 2;
            strips[r] += trips;
        }
        __syncthreads();
        if (tl < m) {
            sperm[r][tl] = newc;
            skofc[r][newc] = tl;
        }
        __syncthreads();
        if (tl < m) {
            if (tl < m - 1) {
                sswap[r][tl][0] += (int)((flags >> tl) & 1u);
                sswap[r][tl][1] += 1;
            }
            const int l = slab[r][newc];
            if (l == 1) shist[r][tl][0] += 1;
            else if (l == 2) shist[r][tl][1] += 1;
            // byte c of sigma: the set-0 configuration at the temperature of set-1 configuration c
            if (r == 1) atomicOr(&sigma[tl >> 2], (uint32_t)sperm[0][skofc[1][tl]] << (8 * (tl & 3)));
        }
        // thresholds of the new temperatures
        eaf_planes(tid, m, skofc, TT, TA, P, ALW);
        for (int e = tid; e < 32 * 32 + 32 * 16; e += EAF_THREADS) M1[e] = 0;
        __syncthreads();
        // overlap patterns (as in ea_measure_ms): row pattern of set-0 configuration c
        // at QP[c], then XORed with the set-1 configuration at the same temperature
        if ((uint32_t)tid < rows) {
            const uint32_t row = tid;
            uint32_t sg[8];
            *(uint4*)sg = *(const uint4*)sigma;
            *(uint4*)(sg + 4) = *(const uint4*)(sigma + 4);
            if constexpr (ROW16) {
                uint32_t v[16];
                eaf_row_words<16>(W0, row, L32, v);
                transpose16x2(v);
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (c < m) QP[c * rowsP + row] = (QT)(c < 16 ? v[c] & 0xFFFFu : v[c - 16] >> 16);
                eaf_row_words<16>(W1, row, L32, v);
                transpose16x2(v);
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (c < m) {
                        const uint32_t pb = c < 16 ? v[c] & 0xFFFFu : v[c - 16] >> 16;
                        QT* q = QP + ((sg[c >> 2] >> (8 * (c & 3))) & 0xFFu) * rowsP + row;
                        *q = (QT)(~((uint32_t)*q ^ pb) & rowmask);
                        if ((c & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                    }
            } else {
                uint32_t v[32];
                eaf_row_words<32>(W0, row, L32, v);
                transpose32(v);
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (c < m) QP[c * rowsP + row] = (QT)v[c];
                eaf_row_words<32>(W1, row, L32, v);
                transpose32(v);
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (c < m) {
                        QT* q = QP + ((sg[c >> 2] >> (8 * (c & 3))) & 0xFFu) * rowsP + row;
                        *q = (QT)(~((uint32_t)*q ^ v[c]) & rowmask);
                        if ((c & 3) == 3) __builtin_amdgcn_sched_barrier(0);
                    }
            }
        }
        __syncthreads();
        if (tid < 8) sigma[tid] = 0u;
        // per temperature, sixteen lanes with a band of rows each: q and the link
        // overlap from popcounts, row sums into M1 (per x1) and M2 (per x2), and
        // per-column counts of q = +1 in bytes (column 4n + b in byte b of cx[n])
        constexpr int NIB = ROW16 ? 4 : 8;
        int q = 0, ql = 0;
        double re0 = 0.0, im0 = 0.0;
        if (mk < m) {
            const QT* qk = QP + sperm[0][mk] * rowsP;
            uint32_t cx[NIB];
#pragma unroll
            for (int n = 0; n < NIB; ++n) cx[n] = 0;
            uint32_t x1 = (uint32_t)mj, x2 = 0;
            while (x1 >= L32) {
                x1 -= L32;
                ++x2;
            }
            for (uint32_t row = mj; row < rows; row += 16) {
                const uint32_t qp = qk[row];
                const int rs = 2 * __popc(qp) - L;
                q += rs;
                const uint32_t rot = ((qp >> 1) | (qp << (L32 - 1))) & rowmask;
                ql += 2 * __popc(~(qp ^ rot) & rowmask) - L;
                const uint32_t ry = (x1 + 1 == L32 ? 0u : x1 + 1) + L32 * x2;
                ql += 2 * __popc(~(qp ^ (uint32_t)qk[ry]) & rowmask) - L;
                atomicAdd(&M1[mk * 32 + x1], rs);
                if (D >= 3) {
                    const uint32_t rz = x1 + L32 * (x2 + 1 == L32 ? 0u : x2 + 1);
                    ql += 2 * __popc(~(qp ^ (uint32_t)qk[rz]) & rowmask) - L;
                    atomicAdd(&M2[mk * 16 + x2], rs);
                }
                // bit b of nibble n to byte b: the four shifted copies of the
                // nibble 7 bits apart never overlap, and bit b of copy b is bit 8 b
#pragma unroll
                for (int n = 0; n < NIB; ++n) cx[n] += __umul24((qp >> (4 * n)) & 15u, 0x204081u) & 0x01010101u;
                x1 += 16;
                while (x1 >= L32) {
                    x1 -= L32;
                    ++x2;
                }
            }
            // this lane's share of sum_x n_x e^(i 2 pi x / L), n_x = sites of column x with q = +1
#pragma unroll
            for (int n = 0; n < NIB; ++n)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int x = 4 * n + b;
                    if (x < L) {
                        const double c = (double)((cx[n] >> (8 * b)) & 0xFFu);
                        re0 += c * ctab[x];
                        im0 += c * stab[x];
                    }
                }
        }
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            q += __shfl_xor(q, off, 64);
            ql += __shfl_xor(ql, off, 64);
            re0 += __shfl_xor(re0, off, 64);
            im0 += __shfl_xor(im0, off, 64);
        }
        __syncthreads();
        // y and z Fourier terms from the row sums, lane mj taking x1 / x2 = mj, mj + 16
        double re1 = 0.0, im1 = 0.0, re2 = 0.0, im2 = 0.0;
        if (mk < m)
            for (int v = mj; v < L; v += 16) {
                re1 += M1[mk * 32 + v] * ctab[v];
                im1 += M1[mk * 32 + v] * stab[v];
                if (D >= 3) {
                    re2 += M2[mk * 16 + v] * ctab[v];
                    im2 += M2[mk * 16 + v] * stab[v];
                }
            }
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            re1 += __shfl_xor(re1, off, 64);
            im1 += __shfl_xor(im1, off, 64);
            if (D >= 3) {
                re2 += __shfl_xor(re2, off, 64);
                im2 += __shfl_xor(im2, off, 64);
            }
        }
        if (mj == 0 && mk < m) {
            const int k = mk;
            const double RE[3] = {2.0 * re0, re1, re2}, IM[3] = {2.0 * im0, im1, im2};
            double qk2 = 0.0;
            for (int ax = 0; ax < D; ++ax) qk2 += (RE[ax] * RE[ax] + IM[ax] * IM[ax]) / ((double)N * (double)N);
            qk2 /= D;
            const double Ea = (double)sE[0][sperm[0][k]];
            const double Eb = (double)sE[1][sperm[1][k]];
            const int bin = 63 - __clzll((long long)(t + 1));
            if (bin != cur_bin) {  // the finished bin goes out, the next one comes in
                if (cur_bin < a.n_bins)
                    for (int o = 0; o < EA_NOBS; ++o)
                        a.acc[acc0 + ((int64_t)cur_bin * m + k) * EA_NOBS + o] = sacc[k][o];
                if (bin < a.n_bins)
                    for (int o = 0; o < EA_NOBS; ++o)
                        sacc[k][o] = a.acc[acc0 + ((int64_t)bin * m + k) * EA_NOBS + o];
                cur_bin = bin;
            }
            if (bin < a.n_bins) {
                double* o = sacc[k];
                const double qq = (double)q / (double)N;
                o[0] += 0.5 * (Ea + Eb) / (double)N;
                o[1] += (double)ql / (double)(D * N);
                o[2] += fabs(qq);
                o[3] += qq * qq;
                o[4] += qq * qq * qq * qq;
                o[5] += qk2;
                o[6] += 1.0;
            }
            if (a.e_series) {
                a.e_series[(((int64_t)s * 2 + 0) * m + k) * a.n_series + t] = (int32_t)Ea;
                a.e_series[(((int64_t)s * 2 + 1) * m + k) * a.n_series + t] = (int32_t)Eb;
            }
            if (a.q_series) a.q_series[((int64_t)s * m + k) * a.n_series + t] = q;
        }
        __syncthreads();  // the next sweep's queues reuse the measurement's LDS
    }
    for (uint32_t t = tid; t < 2 * N; t += EAF_THREADS) gw[t] = dyn[t];
    if (tl < m) {
        a.energies[cfg0 + tl] = sE[r][tl];
        a.sweeps[cfg0 + tl] = sweep;
        a.perm[cfg0 + tl] = sperm[r][tl];
        a.kofc[cfg0 + tl] = skofc[r][tl];
        a.labels[cfg0 + tl] = slab[r][tl];
        a.label_hist[(cfg0 + tl) * 2] += shist[r][tl][0];
        a.label_hist[(cfg0 + tl) * 2 + 1] += shist[r][tl][1];
        if (tl < m - 1) {
            a.swap_acc[((int64_t)s * 2 + r) * (m - 1) + tl] += sswap[r][tl][0];
            a.swap_try[((int64_t)s * 2 + r) * (m - 1) + tl] += sswap[r][tl][1];
        }
    }
    if (tl == 0 && strips[r]) a.trips[s * 2 + r] += strips[r];
    if (mj == 0 && mk < m && cur_bin < a.n_bins)
        for (int o = 0; o < EA_NOBS; ++o) a.acc[acc0 + ((int64_t)cur_bin * m + mk) * EA_NOBS + o] = sacc[mk][o];
}

// E = -sum_b J_b s_i s_j for every configuration (critmc/kernels/glass.py:33-38).
template <int Z, bool G>
__global__ __launch_bounds__(256) void ea_energy(EaArgs a)
{
    __shared__ int red[16];
    __shared__ double redf[16];
    const int m = a.m;
    const int s = blockIdx.x / (2 * m);
    const int rc = blockIdx.x - s * 2 * m;
    const int L = a.L[s];
    const CbGeo g = ea_geo(Z / 2, L);
    const int64_t N = g.N;
    const int8_t* sp = a.spins + a.spin_off[s] + (int64_t)rc * N;
    const int8_t* J = G ? nullptr : a.J + a.j_off[s];
    const double* Jf = G ? a.Jf + a.j_off[s] : nullptr;
    const FastDiv dl = make_fastdiv((uint32_t)L);
    int e = 0;
    double ef = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        int x[3];
        ea_coords<Z / 2>((uint32_t)i, (uint32_t)L, dl, x);
#pragma unroll
        for (int ax = 0; ax < Z / 2; ++ax) {
            const int64_t stride = ax == 0 ? 1 : (ax == 1 ? L : (int64_t)L * L);
            const int64_t j = i + (x[ax] + 1 == L ? -(int64_t)(L - 1) * stride : stride);
            if constexpr (G) ef -= Jf[ax * N + i] * (double)(sp[i] * sp[j]);
            else e -= J[ax * N + i] * sp[i] * sp[j];
        }
    }
    if constexpr (G) {
        ef = wave_sum_f64(ef);
        if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = ef;
    } else {
        e = wave_sum(e);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)((blockDim.x + 63) >> 6);
        if constexpr (G) {
            double t = 0.0;
            for (int w = 0; w < nw; ++w) t += redf[w];
            a.energies_f[(int64_t)s * 2 * m + rc] = t;
        } else {
            int t = 0;
            for (int w = 0; w < nw; ++w) t += red[w];
            a.energies[(int64_t)s * 2 * m + rc] = t;
        }
    }
}

// Random start: site i of configuration (r, c) of sample s takes bit 31 of
// word i&3 of block i>>2 at (sweep 0, colour 0, TAG_INIT, stream r*m + c).
__global__ void ea_init_kernel(EaArgs a)
{
    const int cfg = blockIdx.y;
    const int s = cfg / (2 * a.m), rc = cfg - (cfg / (2 * a.m)) * 2 * a.m;
    const int L = a.L[s];
    const int64_t N = (int64_t)L * (a.d >= 2 ? L : 1) * (a.d >= 3 ? L : 1);
    int8_t* sp = a.spins + a.spin_off[s] + (int64_t)rc * N;
    const uint32_t k0 = a.keys[2 * s], k1 = a.keys[2 * s + 1];
    for (int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; blk * 4 < N;
         blk += (int64_t)gridDim.x * blockDim.x) {
        const u32x4 r = draw(k0, k1, (uint32_t)rc, 0, 0, TAG_INIT, (uint32_t)blk);
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
        for (int k = 0; k < 4 && blk * 4 + k < N; ++k) sp[blk * 4 + k] = (w[k] >> 31) ? 1 : -1;
    }
}

template <bool G>
static void launch_step(const EaArgs& a, int ndim, int sweep_blocks, int block, size_t lds,
                        int swap_threads, uint64_t step, int64_t t, hipStream_t st, bool all_lds)
{
    if (ndim == 2) {
        if (all_lds) hipLaunchKernelGGL((ea_sweep<4, G, true>), dim3(sweep_blocks), dim3(block), lds, st, a);
        else hipLaunchKernelGGL((ea_sweep<4, G, false>), dim3(sweep_blocks), dim3(block), 0, st, a);
        hipLaunchKernelGGL((ea_swap<G>), dim3(swap_threads), dim3(64), 0, st, a, step);
        if (all_lds) hipLaunchKernelGGL((ea_measure<4, G, true>), dim3(a.S * a.m), dim3(256), lds, st, a, t);
        else hipLaunchKernelGGL((ea_measure<4, G, false>), dim3(a.S * a.m), dim3(256), 0, st, a, t);
    } else {
        if (all_lds) hipLaunchKernelGGL((ea_sweep<6, G, true>), dim3(sweep_blocks), dim3(block), lds, st, a);
        else hipLaunchKernelGGL((ea_sweep<6, G, false>), dim3(sweep_blocks), dim3(block), 0, st, a);
        hipLaunchKernelGGL((ea_swap<G>), dim3(swap_threads), dim3(64), 0, st, a, step);
        if (all_lds) hipLaunchKernelGGL((ea_measure<6, G, true>), dim3(a.S * a.m), dim3(256), lds, st, a, t);
        else hipLaunchKernelGGL((ea_measure<6, G, false>), dim3(a.S * a.m), dim3(256), 0, st, a, t);
    }
}

static void launch_energy(const EaArgs& a, int gauss, hipStream_t st)
{
    const int blocks = a.S * 2 * a.m;
    if (a.d == 2) {
        if (gauss) hipLaunchKernelGGL((ea_energy<4, true>), dim3(blocks), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((ea_energy<4, false>), dim3(blocks), dim3(256), 0, st, a);
    } else {
        if (gauss) hipLaunchKernelGGL((ea_energy<6, true>), dim3(blocks), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((ea_energy<6, false>), dim3(blocks), dim3(256), 0, st, a);
    }
}

}  // namespace cm

using namespace cm;

extern "C" {

// Thresholds, then (multi-spin +-J path) per-sample word offsets, the packed
// words and the neighbour-coupling bytes.
static int64_t ea_ws_layout(int32_t n_samples, int32_t m, int64_t total_sites, int64_t* off_ms,
                            int64_t* off_w, int64_t* off_j)
{
    const int64_t thr = ((int64_t)n_samples * m * EA_TAB * 8 + 255) / 256 * 256;
    const int64_t offs = ((int64_t)n_samples * 16 + 255) / 256 * 256;
    const int64_t words = (2 * total_sites * 4 + 255) / 256 * 256;
    if (off_ms) *off_ms = thr;
    if (off_w) *off_w = thr + offs;
    if (off_j) *off_j = thr + offs + words;
    return thr + offs + words + (total_sites + 255) / 256 * 256;
}

int64_t cm_ea_workspace_bytes(int32_t n_samples, int32_t m, int64_t total_sites)
{
    return ea_ws_layout(n_samples, m, total_sites, nullptr, nullptr, nullptr);
}

double cm_det_exp(double x) { return det_exp(x); }

// The multi-spin kernels take +-J batches of at most 32 temperatures whose
// largest lattice keeps a set in LDS and a row in one word (L <= 32, at most
// EA_MS_ROWS rows).
static bool ea_use_ms(int32_t jtype, int32_t m, int32_t ndim, int32_t max_L)
{
    const int64_t maxN = (int64_t)max_L * max_L * (ndim == 3 ? max_L : 1);
    return !jtype && m <= 32 && maxN <= EA_LDS_MAX && max_L <= 32 && maxN / max_L <= EA_MS_ROWS;
}

int cm_ea_multispin(int32_t jtype, int32_t m, int32_t ndim, int32_t max_L)
{
    return ea_use_ms(jtype, m, ndim, max_L) ? 1 : 0;
}

int cm_ea_init(void* spins, const int32_t* L_dev, const int64_t* spin_off_dev, int32_t n_samples,
               int32_t m, int32_t ndim, const uint32_t* keys_dev, void* stream)
{
    CM_REQUIRE(ndim == 2 || ndim == 3, "EA tempering supports square and cubic lattices");
    CM_REQUIRE(n_samples >= 1 && m >= 2 && (int64_t)n_samples * 2 * m <= 65535,
               "need 1..65535 configurations in all");
    EaArgs a{};
    a.spins = (int8_t*)spins;
    a.L = L_dev;
    a.spin_off = spin_off_dev;
    a.keys = keys_dev;
    a.S = n_samples;
    a.m = m;
    a.d = ndim;
    hipLaunchKernelGGL(ea_init_kernel, dim3(8, n_samples * 2 * m), dim3(256), 0,
                       (hipStream_t)stream, a);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_ea_energy(void* spins, const void* J, int32_t jtype, const int32_t* L_dev,
                 const int64_t* spin_off_dev, const int64_t* j_off_dev, int32_t n_samples,
                 int32_t m, int32_t ndim, void* energies_dev, void* stream)
{
    CM_REQUIRE(ndim == 2 || ndim == 3, "EA tempering supports square and cubic lattices");
    CM_REQUIRE(n_samples >= 1 && m >= 2, "need at least one sample and two temperatures");
    CM_REQUIRE(jtype == 0 || jtype == 1, "jtype must be 0 (int8 +-1) or 1 (FP64 couplings)");
    EaArgs a{};
    a.spins = (int8_t*)spins;
    a.J = jtype ? nullptr : (const int8_t*)J;
    a.Jf = jtype ? (const double*)J : nullptr;
    a.L = L_dev;
    a.spin_off = spin_off_dev;
    a.j_off = j_off_dev;
    a.energies = jtype ? nullptr : (int32_t*)energies_dev;
    a.energies_f = jtype ? (double*)energies_dev : nullptr;
    a.S = n_samples;
    a.m = m;
    a.d = ndim;
    launch_energy(a, jtype, (hipStream_t)stream);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_ea_pt_run(void* spins, const void* J, int32_t jtype, const int32_t* L_dev,
                 const int64_t* spin_off_dev, const int64_t* j_off_dev, int32_t n_samples,
                 int32_t m, int32_t ndim, const int32_t* L_host, const double* betas_host,
                 const double* betas_dev, const uint32_t* keys_dev, uint64_t* sweeps_dev,
                 void* energies_dev,
                 int32_t* perm_dev, int32_t* kofc_dev, int32_t* labels_dev, int32_t* trips_dev,
                 int32_t* label_hist_dev, int32_t* swap_acc_dev, int32_t* swap_try_dev,
                 double* acc_dev, int32_t n_bins, void* e_series_dev, int32_t* q_series_dev,
                 int32_t code, int32_t sweeps_per_step, uint64_t step0, int64_t n_steps,
                 void* workspace, int64_t workspace_bytes, int32_t block, void* stream)
{
    CM_REQUIRE(ndim == 2 || ndim == 3, "EA tempering supports square and cubic lattices");
    CM_REQUIRE(n_samples >= 1 && m >= 2, "need at least one sample and two temperatures");
    CM_REQUIRE(code == CODE_METROPOLIS || code == CODE_HEATBATH,
               "update code %d has no checkerboard EA kernel", code);
    CM_REQUIRE(jtype == 0 || jtype == 1, "jtype must be 0 (int8 +-1) or 1 (FP64 couplings)");
    CM_REQUIRE(sweeps_per_step >= 1, "sweeps_per_step must be positive");
    CM_REQUIRE(n_steps >= 0 && n_bins >= 1, "bad step or bin count");
    int max_L = 2;
    for (int s = 0; s < n_samples; ++s) {
        CM_REQUIRE(L_host[s] >= 2 && L_host[s] % 2 == 0 && L_host[s] <= EA_LMAX,
                   "extents must be even, >= 2 and <= 256 (sample %d has %d)", s, L_host[s]);
        max_L = std::max(max_L, (int)L_host[s]);
    }
    int64_t total_sites = 0;
    std::vector<int64_t> ms_off(n_samples);
    for (int s = 0; s < n_samples; ++s) {
        const int64_t Ls = L_host[s];
        ms_off[s] = 2 * total_sites;
        total_sites += Ls * Ls * (ndim == 3 ? Ls : 1);
    }
    CM_REQUIRE(workspace_bytes >= cm_ea_workspace_bytes(n_samples, m, total_sites), "workspace too small");
    if (n_steps == 0) return CM_OK;
    const int z = 2 * ndim;
    std::vector<uint64_t> thr((size_t)n_samples * m * EA_TAB, 0);
    for (int s = 0; s < n_samples; ++s)
        for (int k = 0; k < m; ++k) {
            const double b = betas_host[(size_t)s * m + k];
            CM_REQUIRE(std::isfinite(b), "beta must be finite");
            uint64_t* T = &thr[((size_t)s * m + k) * EA_TAB];
            for (int v = -z; v <= z; ++v) {
                double p;
                if (code == CODE_METROPOLIS) p = v <= 0 ? 1.0 : std::exp(-b * 2.0 * v);
                else p = 1.0 / (1.0 + std::exp(-2.0 * b * v));
                T[v + z] = !(p < 1.0) ? (1ull << 32) : (!(p > 0.0) ? 0 : (uint64_t)std::ceil(p * 4294967296.0));
            }
        }
    hipStream_t st = (hipStream_t)stream;
    uint64_t* dthr = (uint64_t*)workspace;
    CM_HIP_CHECK(upload_async(dthr, thr.data(), thr.size() * 8, st));
    EaArgs a{};
    a.spins = (int8_t*)spins;
    a.J = jtype ? nullptr : (const int8_t*)J;
    a.Jf = jtype ? (const double*)J : nullptr;
    a.L = L_dev;
    a.spin_off = spin_off_dev;
    a.j_off = j_off_dev;
    a.keys = keys_dev;
    a.betas = betas_dev;
    a.thr = dthr;
    a.sweeps = sweeps_dev;
    a.energies = jtype ? nullptr : (int32_t*)energies_dev;
    a.energies_f = jtype ? (double*)energies_dev : nullptr;
    a.perm = perm_dev;
    a.kofc = kofc_dev;
    a.labels = labels_dev;
    a.trips = trips_dev;
    a.label_hist = label_hist_dev;
    a.swap_acc = swap_acc_dev;
    a.swap_try = swap_try_dev;
    a.acc = acc_dev;
    a.e_series = jtype ? nullptr : (int32_t*)e_series_dev;
    a.e_series_f = jtype ? (double*)e_series_dev : nullptr;
    a.q_series = q_series_dev;
    a.S = n_samples;
    a.m = m;
    a.d = ndim;
    a.code = code;
    a.sweeps_per_step = sweeps_per_step;
    a.n_bins = n_bins;
    a.n_series = n_steps;
    if (block <= 0) block = 64;
    CM_REQUIRE(block % 64 == 0 && block <= 256, "block must be a multiple of 64, <= 256");
    const int64_t maxN = (int64_t)max_L * max_L * (ndim == 3 ? max_L : 1);
    const bool all_lds = maxN <= EA_LDS_MAX;
    const size_t lds = all_lds ? (size_t)maxN : 0;
    const int sweep_blocks = n_samples * 2 * m;
    const int swap_threads = 2 * n_samples;
    if (ea_use_ms(jtype, m, ndim, max_L)) {
        // multi-spin path: pack, step, unpack
        int64_t o_ms, o_w, o_j;
        ea_ws_layout(n_samples, m, total_sites, &o_ms, &o_w, &o_j);
        // dispatch the largest sets first, so the heavy workgroups spread over the CUs
        std::vector<int64_t> hdr(2 * (size_t)n_samples, 0);
        std::copy(ms_off.begin(), ms_off.end(), hdr.begin());
        int32_t* order = (int32_t*)(hdr.data() + n_samples);
        const bool fused = g_ea_fused != 0;
        if (fused) {
            // one workgroup per sample, largest first: a CU that finishes a sample
            // takes the largest one still waiting (longest-first list scheduling).
            // Order 1: the larger half largest first, then the smaller half
            // smallest first, pairing workgroups b and b + ceil(S/2).
            std::vector<int32_t> by(n_samples);
            for (int v = 0; v < n_samples; ++v) by[v] = v;
            std::stable_sort(by.begin(), by.end(), [&](int32_t x, int32_t y) { return L_host[x] > L_host[y]; });
            const int H = (n_samples + 1) / 2;
            for (int b = 0; b < n_samples; ++b)
                order[b] = g_ea_order == 1 && b >= H ? by[n_samples - 1 - (b - H)] : by[b];
        } else {
            for (int v = 0; v < 2 * n_samples; ++v) order[v] = v;
            std::stable_sort(order, order + 2 * n_samples,
                             [&](int32_t x, int32_t y) { return L_host[x >> 1] > L_host[y >> 1]; });
        }
        int64_t* d_off = (int64_t*)((char*)workspace + o_ms);
        CM_HIP_CHECK(upload_async(d_off, hdr.data(), hdr.size() * 8, st));
        a.ms_off = d_off;
        a.ms_order = (const int32_t*)(d_off + n_samples);
        a.samples = a.ms_order;
        a.msw = (uint32_t*)((char*)workspace + o_w);
        a.jneg = (uint8_t*)((char*)workspace + o_j);
        const dim3 pgrid((unsigned)std::min<int64_t>((maxN + 255) / 256, 256), (unsigned)n_samples);
        if (ndim == 2) hipLaunchKernelGGL(ea_ms_pack<2>, pgrid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL(ea_ms_pack<3>, pgrid, dim3(256), 0, st, a);
        if (fused) {
            const bool row16 = max_L <= 16;
            const size_t dyn = (size_t)(8 * maxN + (maxN + 15) / 16 * 16 +
                                        std::max<int64_t>(EAF_UNION, eaf_union_need(row16)));
            const void* fn = ndim == 2 ? (row16 ? (const void*)ea_ms_fused<4, true> : (const void*)ea_ms_fused<4, false>)
                                       : (const void*)ea_ms_fused<6, true>;
            CM_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
            if (ndim == 2 && row16)
                hipLaunchKernelGGL((ea_ms_fused<4, true>), dim3(n_samples), dim3(EAF_THREADS), dyn, st, a, step0, n_steps);
            else if (ndim == 2)
                hipLaunchKernelGGL((ea_ms_fused<4, false>), dim3(n_samples), dim3(EAF_THREADS), dyn, st, a, step0, n_steps);
            else
                hipLaunchKernelGGL((ea_ms_fused<6, true>), dim3(n_samples), dim3(EAF_THREADS), dyn, st, a, step0, n_steps);
            hipLaunchKernelGGL(ea_ms_unpack, pgrid, dim3(256), 0, st, a);
            CM_HIP_CHECK(hipGetLastError());
            return CM_OK;
        }
        const bool ms_lds = maxN * 5 <= 60 * 1024;
        const size_t mlds = ms_lds ? (size_t)maxN * 5 : 0;
        for (int64_t t = 0; t < n_steps; ++t) {
            const uint64_t step = step0 + (uint64_t)t;
            if (ndim == 2) {
                if (ms_lds) hipLaunchKernelGGL((ea_ms_sweep<4, true>), dim3(2 * n_samples), dim3(EA_MS_THREADS), mlds, st, a);
                else hipLaunchKernelGGL((ea_ms_sweep<4, false>), dim3(2 * n_samples), dim3(EA_MS_THREADS), 0, st, a);
                hipLaunchKernelGGL((ea_swap<false>), dim3(swap_threads), dim3(64), 0, st, a, step);
                if (max_L <= 16) hipLaunchKernelGGL((ea_measure_ms<4, true>), dim3(a.S), dim3(256), 0, st, a, t);
                else hipLaunchKernelGGL((ea_measure_ms<4, false>), dim3(a.S), dim3(256), 0, st, a, t);
            } else {
                if (ms_lds) hipLaunchKernelGGL((ea_ms_sweep<6, true>), dim3(2 * n_samples), dim3(EA_MS_THREADS), mlds, st, a);
                else hipLaunchKernelGGL((ea_ms_sweep<6, false>), dim3(2 * n_samples), dim3(EA_MS_THREADS), 0, st, a);
                hipLaunchKernelGGL((ea_swap<false>), dim3(swap_threads), dim3(64), 0, st, a, step);
                hipLaunchKernelGGL((ea_measure_ms<6, true>), dim3(a.S), dim3(256), 0, st, a, t);
            }
        }
        hipLaunchKernelGGL(ea_ms_unpack, pgrid, dim3(256), 0, st, a);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    }
    int64_t next_edge = 1;
    int bin = 0;
    for (int64_t t = 0; t < n_steps; ++t) {
        const uint64_t step = step0 + (uint64_t)t;
        if (t >= next_edge) {  // bin edge: FP64 energies are recomputed (integer ones are exact)
            ++bin;
            next_edge = ((int64_t)1 << (bin + 1)) - 1;
            if (jtype) launch_energy(a, 1, st);
        }
        if (jtype) launch_step<true>(a, ndim, sweep_blocks, block, lds, swap_threads, step, t, st, all_lds);
        else launch_step<false>(a, ndim, sweep_blocks, block, lds, swap_threads, step, t, st, all_lds);
    }
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // extern "C"
