// Checkerboard Ising sweeps for gfx950: Metropolis and heat bath, int8 or
// 32-spin bit-packed state.
//
// Reference semantics (critmc/kernels/ising.py):
//   Metropolis  :47-56 table, :68-87 sweep   a = s_i sum_j s_j; accept iff p >= 1 or u < p
//   heat bath   :59-65 table, :90-107 sweep  new = +1 iff u < 1/(1+exp(-2 beta (f+h)))
//   run         :173-237 per measurement E_b, M, accepted
// Visiting order is colour 0 then colour 1 (colour = sum of coordinates mod 2).
//
// Layout.  The colour-c sites of a lattice row r (x1.. fixed) are
// x0 = 2m + ((x1 + x2 + c) & 1), m < M = L0/2, packed 32 per word:
// group g = r * G + m / 32 (G = ceil(M/32)), bit m % 32, bit set = spin +1.
// The packed HBM image is [replica][colour][row][G] u32, which is also the
// LDS image of the batched kernel.  Horizontal neighbours of a word are the
// other colour's word at the same position and the same word shifted by one
// bit (carry from the adjacent word); vertical neighbours are the other
// colour's words of the adjacent rows at the same position.
//
// Uniforms.  A group's 32 sites share 32 Philox words W_0..W_31 (8 blocks,
// tag TAG_ISING); the site at bit b has x = sum_j bit_b(W_j) 2^(31-j), and
// "u < p" is "x < T" with T = ceil(p 2^32) computed on the host in FP64.  The
// comparison runs plane by plane (most significant first) on all 32 lanes of
// a word at once and stops as soon as every lane is decided -- usually after
// one or two Philox blocks instead of eight.  The early stop does not change
// any decision, so the result equals the full 32-bit comparison.
//
// Forms: plain (one chain, one launch per half sweep), batched (one launch,
// replicas resident in LDS; undecided words leave their wave for an LDS queue)
// and wide (one launch per half sweep over all replicas, state in HBM: 2D
// rows of 64-site segments, or any lattice through the plain form's words).
// All draw the same uniforms for the same word, so they run the same chains.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "cb_geom.hpp"
#include "common.hpp"
#include "philox.hpp"
#include "ising_core.hpp"

namespace cm {

// ---------------------------------------------------------------------------
// Plain form: one chain, one launch per half sweep, int8 spins in HBM.

struct Geo {
    int d, L0, L1, L2, M, G;
    int64_t rows;
};

// One word of the half sweep: word t = r G + w of colour `colour` (row r of
// the lattice's rows, 32-site group w), its spins gathered from the int8
// lattice s, updated and scattered back.  The group index t selects the
// word's Philox blocks, so every form that calls this draws the same
// uniforms.
template <int Z, int NCMAX>
__device__ __forceinline__ void plain_word(int8_t* __restrict__ s, const Geo& g, int code,
                                           int colour, uint64_t sweep, uint32_t k0, uint32_t k1,
                                           uint32_t sid, const uint32_t* __restrict__ tabs,
                                           int64_t t, Delta& d)
{
    const int64_t r = t / g.G;
    const int w = (int)(t - r * g.G);
    int x1 = 0, x2 = 0;
    if (Z >= 4) {
        x1 = (int)(r % g.L1);
        x2 = (int)(r / g.L1);
    }
    const int p = (x1 + x2 + colour) & 1;
    const int64_t base = r * g.L0;
    int64_t rows_nb[4] = {0, 0, 0, 0};
    if (Z >= 4) {
        const int64_t y0 = r - x1;
        rows_nb[0] = (y0 + (x1 + 1 == g.L1 ? 0 : x1 + 1)) * g.L0;
        rows_nb[1] = (y0 + (x1 == 0 ? g.L1 - 1 : x1 - 1)) * g.L0;
    }
    if (Z == 6) {
        const int64_t plane = (int64_t)g.L1;
        const int64_t z0 = r - (int64_t)x2 * plane;
        rows_nb[2] = (z0 + (int64_t)(x2 + 1 == g.L2 ? 0 : x2 + 1) * plane) * g.L0;
        rows_nb[3] = (z0 + (int64_t)(x2 == 0 ? g.L2 - 1 : x2 - 1) * plane) * g.L0;
    }
    const int nbits = min(32, g.M - 32 * w);
    uint32_t own = 0;
    uint32_t nb[Z];
#pragma unroll
    for (int k = 0; k < Z; ++k) nb[k] = 0;
    for (int b = 0; b < nbits; ++b) {
        const int x0 = 2 * (32 * w + b) + p;
        const int xp = x0 + 1 == g.L0 ? 0 : x0 + 1;
        const int xm = x0 == 0 ? g.L0 - 1 : x0 - 1;
        own |= (uint32_t)(s[base + x0] > 0) << b;
        nb[0] |= (uint32_t)(s[base + xp] > 0) << b;
        nb[1] |= (uint32_t)(s[base + xm] > 0) << b;
#pragma unroll
        for (int k = 2; k < Z; ++k) nb[k] |= (uint32_t)(s[rows_nb[k - 2] + x0] > 0) << b;
    }
    const uint32_t valid = nbits == 32 ? ~0u : ((1u << nbits) - 1u);
    const uint32_t nw = update_word<Z, NCMAX>(code, own, nb, valid, tabs, k0, k1, sid, sweep,
                                              (uint32_t)colour, (uint32_t)t, d);
    uint32_t changed = (nw ^ own) & valid;
    while (changed) {
        const int b = __ffs(changed) - 1;
        changed &= changed - 1;
        s[base + 2 * (32 * w + b) + p] = ((nw >> b) & 1u) ? (int8_t)1 : (int8_t)-1;
    }
}

template <int Z, int NCMAX>
__global__ __launch_bounds__(256) void ising_plain_half(int8_t* __restrict__ s, Geo g, int code,
                                                        int colour, uint64_t sweep, uint32_t k0,
                                                        uint32_t k1, uint32_t sid,
                                                        const uint32_t* __restrict__ tabs,
                                                        int32_t* __restrict__ de_out,
                                                        int32_t* __restrict__ dm_out,
                                                        int32_t* __restrict__ acc_out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Delta d{0, 0, 0};
    if (t < g.rows * g.G) plain_word<Z, NCMAX>(s, g, code, colour, sweep, k0, k1, sid, tabs, t, d);
    const int de = wave_sum(d.de), dm = wave_sum(d.dm), ac = wave_sum(d.acc);
    if ((threadIdx.x & 63) == 0 && (de | dm | ac)) {
        atomicAdd(de_out, de);
        atomicAdd(dm_out, dm);
        atomicAdd(acc_out, ac);
    }
}

// Wide form on any lattice the plain kernel sweeps (chains, 2D with any even
// L0, cubic): one launch per half sweep over (word chunks x replicas), each
// replica's keys, stream id and sweep counter read on the device, its
// deltas added to its own measurement slot.  The words and uniforms are the
// plain form's, so both give the same chains.
template <int Z, int NCMAX>
__global__ __launch_bounds__(256) void ising_any_half(
    int8_t* __restrict__ spins, Geo g, int code, int colour, uint64_t sweep_add,
    const uint32_t* __restrict__ tabs, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ sids, const uint64_t* __restrict__ sweeps, int64_t n_meas,
    int64_t meas, const int64_t* __restrict__ n_meas_rep, int32_t* __restrict__ de_out,
    int32_t* __restrict__ dm_out, int32_t* __restrict__ acc_out)
{
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x);
    const uint32_t bx = lb - (uint32_t)rep * gridDim.x;
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;  // whole workgroup
    const int64_t t = (int64_t)bx * blockDim.x + threadIdx.x;
    Delta d{0, 0, 0};
    if (t < g.rows * g.G)
        plain_word<Z, NCMAX>(spins + (int64_t)rep * g.rows * g.L0, g, code, colour,
                             sweeps[rep] + sweep_add, keys[2 * rep], keys[2 * rep + 1], sids[rep],
                             tabs + (int64_t)TAB_WORDS * rep, t, d);
    const int de = wave_sum(d.de), dm = wave_sum(d.dm), ac = wave_sum(d.acc);
    if ((threadIdx.x & 63) == 0 && (de | dm | ac)) {
        const int64_t o = (int64_t)rep * n_meas + meas;
        atomicAdd(&de_out[o], de);
        atomicAdd(&dm_out[o], dm);
        atomicAdd(&acc_out[o], ac);
    }
}

__global__ void fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Turns per-measurement deltas into the running E_b, M series (one thread per
// chain).  n_meas_rep: null or per-chain caps; sweeps: null, or advanced by
// per_meas sweeps per recorded measurement.
__global__ void prefix_series(int32_t* __restrict__ e_out, int32_t* __restrict__ m_out,
                              int64_t n_meas, int32_t* __restrict__ e_run,
                              int32_t* __restrict__ m_run, int R,
                              const int64_t* __restrict__ n_meas_rep,
                              uint64_t* __restrict__ sweeps, uint64_t per_meas)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= R) return;
    const int64_t n = n_meas_rep ? min(n_meas, n_meas_rep[rep]) : n_meas;
    int32_t e = e_run[rep], m = m_run[rep];
    for (int64_t k = 0; k < n; ++k) {
        e += e_out[(int64_t)rep * n_meas + k];
        m += m_out[(int64_t)rep * n_meas + k];
        e_out[(int64_t)rep * n_meas + k] = e;
        m_out[(int64_t)rep * n_meas + k] = m;
    }
    e_run[rep] = e;
    m_run[rep] = m;
    if (sweeps) sweeps[rep] += per_meas * (uint64_t)n;
}

// ---------------------------------------------------------------------------
// Batched form: one workgroup per replica (grid-stride), the replica's
// bit-packed state resident in LDS for the whole call, per-measurement
// E_b / M / accepted reduced in the workgroup.  2D square lattices with
// M <= 32 or M a multiple of 32 (L0 <= 64 or L0 a multiple of 64).

__device__ __forceinline__ void unpack_pair_u32(uint32_t v, uint32_t& even, uint32_t& odd, int sh)
{
    // bytes b0..b3 of v are spins (+1 = 0x01, -1 = 0xFF); bit set = +1
    const uint32_t up = ~v;
    even |= (((up >> 7) & 1u) | ((up >> 22) & 2u)) << sh;
    odd |= (((up >> 15) & 1u) | ((up >> 30) & 2u)) << sh;
}

__device__ __forceinline__ uint32_t pack_pair_u32(uint32_t even, uint32_t odd, int sh)
{
    const uint32_t e = even >> sh, o = odd >> sh;
    const uint32_t b0 = (e & 1u) ? 0x01u : 0xFFu;
    const uint32_t b1 = (o & 1u) ? 0x01u : 0xFFu;
    const uint32_t b2 = (e & 2u) ? 0x01u : 0xFFu;
    const uint32_t b3 = (o & 2u) ? 0x01u : 0xFFu;
    return b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
}

template <int LAYOUT>
__device__ void load_replica(const void* __restrict__ spins, int rep, int L0, int L1, int G,
                             uint32_t* __restrict__ lds)
{
    const int W = L1 * G;
    if (LAYOUT == 1) {
        const uint32_t* src = (const uint32_t*)spins + (int64_t)rep * 2 * W;
        for (int i = threadIdx.x; i < 2 * W; i += blockDim.x) lds[i] = src[i];
        return;
    }
    const int8_t* src = (const int8_t*)spins + (int64_t)rep * L0 * L1;
    for (int idx = threadIdx.x; idx < W; idx += blockDim.x) {
        const int y = idx / G, w = idx - (idx / G) * G;
        const int8_t* row = src + (int64_t)y * L0 + 64 * w;
        uint32_t even = 0, odd = 0;
        if ((L0 & 63) == 0) {
            const uint4* v = (const uint4*)row;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 q = v[k];
                unpack_pair_u32(q.x, even, odd, 8 * k + 0);
                unpack_pair_u32(q.y, even, odd, 8 * k + 2);
                unpack_pair_u32(q.z, even, odd, 8 * k + 4);
                unpack_pair_u32(q.w, even, odd, 8 * k + 6);
            }
        } else {
            for (int x = 0; x < L0; ++x) {
                const uint32_t bit = row[x] > 0 ? 1u : 0u;
                if (x & 1) odd |= bit << (x >> 1);
                else even |= bit << (x >> 1);
            }
        }
        const int ce = y & 1;  // colour of even x in row y
        lds[ce * W + idx] = even;
        lds[(1 - ce) * W + idx] = odd;
    }
}

template <int LAYOUT>
__device__ void store_replica(void* __restrict__ spins, int rep, int L0, int L1, int G,
                              const uint32_t* __restrict__ lds)
{
    const int W = L1 * G;
    if (LAYOUT == 1) {
        uint32_t* dst = (uint32_t*)spins + (int64_t)rep * 2 * W;
        for (int i = threadIdx.x; i < 2 * W; i += blockDim.x) dst[i] = lds[i];
        return;
    }
    int8_t* dst = (int8_t*)spins + (int64_t)rep * L0 * L1;
    for (int idx = threadIdx.x; idx < W; idx += blockDim.x) {
        const int y = idx / G, w = idx - (idx / G) * G;
        int8_t* row = dst + (int64_t)y * L0 + 64 * w;
        const int ce = y & 1;
        const uint32_t even = lds[ce * W + idx], odd = lds[(1 - ce) * W + idx];
        if ((L0 & 63) == 0) {
            uint4* v = (uint4*)row;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint4 q;
                q.x = pack_pair_u32(even, odd, 8 * k + 0);
                q.y = pack_pair_u32(even, odd, 8 * k + 2);
                q.z = pack_pair_u32(even, odd, 8 * k + 4);
                q.w = pack_pair_u32(even, odd, 8 * k + 6);
                v[k] = q;
            }
        } else {
            for (int x = 0; x < L0; ++x) {
                const uint32_t bit = ((x & 1) ? (odd >> (x >> 1)) : (even >> (x >> 1))) & 1u;
                row[x] = bit ? (int8_t)1 : (int8_t)-1;
            }
        }
    }
}

// Words of one colour of an LDS-resident replica: own word idx and its four
// neighbour words (same position in the other colour, shifted by one bit
// with the carry from the adjacent word, rows above and below).
__device__ __forceinline__ void batched_word(const uint32_t* __restrict__ Sc,
                                             const uint32_t* __restrict__ So, int idx, int y, int w,
                                             int G, int W, int M, int L1, int c, uint32_t valid,
                                             uint32_t& own, uint32_t (&nb)[4])
{
    const int p = (y + c) & 1;
    const int row = idx - w;  // y * G
    const int rp = y + 1 == L1 ? 0 : row + G;
    const int rm = y == 0 ? W - G : row - G;
    own = Sc[idx];
    const uint32_t same = So[idx];
    uint32_t sh;
    if (M >= 32) {
        // the word before (p = 0) or after (p = 1) supplies the bit shifted in;
        // lanes of both parities share a wave, so no branch
        const int wn = p == 0 ? (w == 0 ? G - 1 : w - 1) : (w == G - 1 ? 0 : w + 1);
        const uint32_t o = So[row + wn];
        sh = p == 0 ? (same << 1) | (o >> 31) : (same >> 1) | (o << 31);
    } else if (p == 0) {
        sh = ((same << 1) | (same >> (M - 1))) & valid;
    } else {
        sh = ((same >> 1) | ((same & 1u) << (M - 1))) & valid;
    }
    nb[0] = same;
    nb[1] = sh;
    nb[2] = So[rp + w];
    nb[3] = So[rm + w];
}

// Philox blocks a word may draw before it leaves the wave's lockstep.  A wave
// keeps drawing until its last lane is decided, about 3.3 blocks for 64 words
// of near-critical Metropolis, while a word needs 1.9 on average.  Words
// still undecided after this budget go to an LDS queue that the whole
// workgroup then drains densely, continuing the same comparison, so every
// decision is unchanged.
#ifndef CM_ISING_DRAW_BUDGET
#define CM_ISING_DRAW_BUDGET 2
#endif
constexpr uint32_t ISING_DRAW_BUDGET = CM_ISING_DRAW_BUDGET;
constexpr int ISING_QCAP = 512;

// ZF: every row of the schedule is Metropolis and every replica's table is the
// zero-field z = 4 one (classes 3 and 4), checked on the host; only that rule
// is compiled, which keeps the kernel's register use down.
template <int LAYOUT, int NCMAX, bool ZF>
__global__ __launch_bounds__(256) void ising2d_batched(
    void* __restrict__ spins, int R, int L0, int L1, const uint32_t* __restrict__ tabs,
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ sids,
    uint64_t* __restrict__ sweeps, Sched sched, int64_t n_meas, int32_t* __restrict__ e_run,
    int32_t* __restrict__ m_run, int32_t* __restrict__ e_out, int32_t* __restrict__ m_out,
    int32_t* __restrict__ acc_out, const int64_t* __restrict__ n_meas_rep)
{
    extern __shared__ uint32_t lds[];
    const int M = L0 >> 1;
    const int G = (M + 31) >> 5;
    const int W = L1 * G;
    int* red = (int*)(lds + 2 * W);         // [16][3]
    uint32_t* queue = lds + 2 * W + 48;     // [3][ISING_QCAP]: words, pending, below
    int* qcount = (int*)(queue + 3 * ISING_QCAP);  // [2], alternating half sweeps
    const int nwaves = (blockDim.x + 63) >> 6;
    const uint32_t valid = M >= 32 ? ~0u : ((1u << M) - 1u);
    const FastDiv dG = make_fastdiv((uint32_t)G);
    const int dy = (int)blockDim.x % G == 0 ? (int)blockDim.x / G : 0;  // rows per stride, 0: divide
    // Row-parity split (rows of whole 32-site words, whole waves and whole
    // rows per half workgroup): the first half of the workgroup takes
    // the even rows, the second half the odd rows.  A wave then holds rows of
    // one parity, so which side supplies the shifted-in bit is uniform across
    // it, and a lane keeps its word column for the whole half sweep.
    const int hb = (int)blockDim.x >> 1;
    const bool rowpar = M >= 32 && (hb & 63) == 0 && hb % G == 0;
    const int par = (int)threadIdx.x >= hb ? 1 : 0;
    const int u = (int)threadIdx.x - par * hb;
    const int u_row = (int)fdiv((uint32_t)u, dG), u_w = u - u_row * G;
    const int y0 = 2 * u_row + par, ystep = rowpar ? 2 * (hb / G) : 1, istep = ystep * G;
    const int hoff0 = u_w == 0 ? G - 1 : -1;     // colour parity 0: the word before, wrapping
    const int hoff1 = u_w == G - 1 ? 1 - G : 1;  // colour parity 1: the word after, wrapping
    if (threadIdx.x < 2) qcount[threadIdx.x] = 0;
    uint32_t hs = 0;

    for (int rep = blockIdx.x; rep < R; rep += gridDim.x) {
        load_replica<LAYOUT>(spins, rep, L0, L1, G, lds);
        __syncthreads();
        const uint32_t k0 = keys[2 * rep], k1 = keys[2 * rep + 1], sid = sids[rep];
        const uint32_t* tab = tabs + (int64_t)TAB_WORDS * rep;
        uint64_t sweep = sweeps[rep];
        int32_t e = e_run[rep], m = m_run[rep];
        const int64_t n_rep_meas = n_meas_rep ? min(n_meas, n_meas_rep[rep]) : n_meas;

        for (int64_t meas = 0; meas < n_rep_meas; ++meas) {
            Delta d{0, 0, 0};
            FlipCounts fc{0, 0, 0, 0, 0};
            for (int row = 0; row < sched.n; ++row) {
                const int code = sched.code[row];
                ClassTab<NCMAX> ct;
                decode_classes<NCMAX>(code == CODE_METROPOLIS ? tab : tab + 32, ct);
                const bool met34 = code == CODE_METROPOLIS && table_is_met34(tab);
                for (int rp = 0; rp < sched.count[row]; ++rp) {
                    for (int c = 0; c < 2; ++c, ++hs) {
                        uint32_t* Sc = lds + c * W;
                        const uint32_t* So = lds + (1 - c) * W;
                        int* qc = qcount + (hs & 1);
                        // one half sweep, the class rule as a template argument:
                        // 0 any table, 1 Metropolis in a field (spin selection),
                        // 2 zero-field Metropolis at z = 4 (classes 3 and 4)
                        auto half = [&](auto mode_tag) {
                        constexpr int MODE = decltype(mode_tag)::value;
                        auto classify = [&](uint32_t own, const uint32_t (&nb)[4],
                                            WordDecision<NCMAX>& wd) {
                            if constexpr (MODE == 2) word_classify_met34<NCMAX>(own, nb, valid, ct, wd);
                            else word_classify_decoded<4, NCMAX, MODE == 1>(code, own, nb, valid, ct, wd);
                        };
                        // decide one word within the draw budget; still undecided, it
                        // goes to the queue (or, with the queue full, finishes here)
                        auto update = [&](int idx, uint32_t own, const uint32_t (&nb)[4]) {
                            WordDecision<NCMAX> wd;
                            classify(own, nb, wd);
                            word_draw<NCMAX>(wd, 0, ISING_DRAW_BUDGET, k0, k1, sid, sweep,
                                             (uint32_t)c, (uint32_t)idx);
                            if (wd.pending) {
                                const int slot = atomicAdd(qc, 1);
                                if (slot < ISING_QCAP) {
                                    queue[slot] = (uint32_t)idx;
                                    queue[ISING_QCAP + slot] = wd.pending;
                                    queue[2 * ISING_QCAP + slot] = wd.below;
                                    return;
                                }
                                word_draw<NCMAX>(wd, ISING_DRAW_BUDGET, 8, k0, k1, sid, sweep,
                                                 (uint32_t)c, (uint32_t)idx);
                            }
                            if (MODE == 2 || code == CODE_METROPOLIS) Sc[idx] = word_apply_metropolis<NCMAX>(own, valid, wd, fc);
                            else Sc[idx] = word_apply<4, NCMAX>(code, own, valid, wd, d);
                        };
                        if (rowpar) {
                            // even rows can wrap only downwards (row 0), odd rows only
                            // upwards (row L1 - 1); the count is symmetric in the
                            // neighbours, so their order does not matter
                            const int wpar = __builtin_amdgcn_readfirstlane(par);
                            const int p = (wpar + c) & 1;
                            const int hoff = p ? hoff1 : hoff0;
                            const int ywrap = wpar ? L1 - 1 : 0;
                            const int off_in = wpar ? -G : G, off_out = -off_in;
                            const int off_wrap = wpar ? G - W : W - G;
                            for (int y = y0, idx = y0 * G + u_w; y < L1; y += ystep, idx += istep) {
                                const uint32_t* q = So + idx;
                                const uint32_t own = Sc[idx], same = q[0], o = q[hoff];
                                uint32_t nb[4];
                                nb[0] = same;
                                nb[1] = p == 0 ? (same << 1) | (o >> 31) : (same >> 1) | (o << 31);
                                nb[2] = q[off_in];
                                nb[3] = q[y == ywrap ? off_wrap : off_out];
                                update(idx, own, nb);
                            }
                        } else {
                            // rows step by blockDim / G when G divides blockDim: no division per word
                            int y = (int)fdiv(threadIdx.x, dG), w = (int)threadIdx.x - y * G;
                            for (int idx = threadIdx.x; idx < W; idx += blockDim.x) {
                                if (dy == 0) {
                                    y = (int)fdiv((uint32_t)idx, dG);
                                    w = idx - y * G;
                                }
                                uint32_t own, nb[4];
                                batched_word(Sc, So, idx, y, w, G, W, M, L1, c, valid, own, nb);
                                y += dy;
                                update(idx, own, nb);
                            }
                        }
                        __syncthreads();
                        const int nq = min(*qc, ISING_QCAP);
                        if (threadIdx.x == 0) qcount[(hs + 1) & 1] = 0;
                        for (int t = threadIdx.x; t < nq; t += blockDim.x) {
                            const int idx = (int)queue[t];
                            const int qy = (int)fdiv((uint32_t)idx, dG);
                            uint32_t own, nb[4];
                            batched_word(Sc, So, idx, qy, idx - qy * G, G, W, M, L1, c, valid, own, nb);
                            WordDecision<NCMAX> wd;
                            classify(own, nb, wd);
                            wd.pending = queue[ISING_QCAP + t];
                            wd.below = queue[2 * ISING_QCAP + t];
                            word_draw<NCMAX>(wd, ISING_DRAW_BUDGET, 8, k0, k1, sid, sweep,
                                             (uint32_t)c, (uint32_t)idx);
                            if (MODE == 2 || code == CODE_METROPOLIS) Sc[idx] = word_apply_metropolis<NCMAX>(own, valid, wd, fc);
                            else Sc[idx] = word_apply<4, NCMAX>(code, own, valid, wd, d);
                        }
                        __syncthreads();
                        };
                        if constexpr (ZF) {
                            half(std::integral_constant<int, 2>{});
                        } else if (ct.anysel) {
                            half(std::integral_constant<int, 1>{});
                        } else {
                            bool done = false;
                            if constexpr (NCMAX == 2) {
                                if (met34) {
                                    half(std::integral_constant<int, 2>{});
                                    done = true;
                                }
                            }
                            if (!done) half(std::integral_constant<int, 0>{});
                        }
                    }
                    ++sweep;
                }
            }
            fold_counts<4>(fc, d);
            const int de = wave_sum_dpp(d.de), dm = wave_sum_dpp(d.dm), ac = wave_sum_dpp(d.acc);
            if ((threadIdx.x & 63) == 0) {
                red[3 * (threadIdx.x >> 6) + 0] = de;
                red[3 * (threadIdx.x >> 6) + 1] = dm;
                red[3 * (threadIdx.x >> 6) + 2] = ac;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int sde = 0, sdm = 0, sac = 0;
                for (int k = 0; k < nwaves; ++k) {
                    sde += red[3 * k];
                    sdm += red[3 * k + 1];
                    sac += red[3 * k + 2];
                }
                e += sde;
                m += sdm;
                const int64_t o = (int64_t)rep * n_meas + meas;
                e_out[o] = e;
                m_out[o] = m;
                acc_out[o] = sac;
            }
            __syncthreads();
        }
        store_replica<LAYOUT>(spins, rep, L0, L1, G, lds);
        if (threadIdx.x == 0) {
            e_run[rep] = e;
            m_run[rep] = m;
            sweeps[rep] = sweep;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Wide form: one launch per half sweep over every replica, one lane per word
// of 32 same-colour sites, state in HBM (int8 or packed).  For a few large
// lattices, where one workgroup per replica would leave most CUs idle.  2D,
// L0 a multiple of 64, so a word covers one 64-byte segment of a row.  Lane
// idx = y * G + w draws the same uniforms as word idx of the batched form, so
// the two forms give identical chains.  A lane rewrites its whole 64-byte
// segment when a spin changed: the other colour's bytes in it are rewritten
// with the values every reader of this launch sees, since that colour is not
// updated in this half sweep.

// Parity-p word (bit m = site x = 2m + p) of the 64 int8 spins at row + 64 w.
__device__ __forceinline__ void load_segment(const int8_t* __restrict__ seg, uint32_t& even,
                                             uint32_t& odd)
{
    const uint4* v = (const uint4*)seg;
    even = 0;
    odd = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 q = v[k];
        unpack_pair_u32(q.x, even, odd, 8 * k + 0);
        unpack_pair_u32(q.y, even, odd, 8 * k + 2);
        unpack_pair_u32(q.z, even, odd, 8 * k + 4);
        unpack_pair_u32(q.w, even, odd, 8 * k + 6);
    }
}

constexpr int WIDE_LANES = 256;  // lanes per workgroup of the wide Ising kernels

// The workgroup's energy, magnetisation and acceptance deltas added to
// measurement slot o with one atomic each: the wave sums are combined in LDS
// first, so a launch issues a quarter of the atomics one per wave would.
__device__ __forceinline__ void wide_deltas(const Delta& d, int32_t* __restrict__ de_out,
                                            int32_t* __restrict__ dm_out,
                                            int32_t* __restrict__ acc_out, int64_t o)
{
    __shared__ int red[3][WIDE_LANES / 64];
    const int de = wave_sum(d.de), dm = wave_sum(d.dm), ac = wave_sum(d.acc);
    const int w = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = de;
        red[1][w] = dm;
        red[2][w] = ac;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t0 = 0, t1 = 0, t2 = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            t0 += red[0][k];
            t1 += red[1][k];
            t2 += red[2][k];
        }
        if (t0 | t1 | t2) {
            atomicAdd(&de_out[o], t0);
            atomicAdd(&dm_out[o], t1);
            atomicAdd(&acc_out[o], t2);
        }
    }
}

template <int LAYOUT, int NC>
__global__ __launch_bounds__(WIDE_LANES) void ising2d_half_wide(
    void* __restrict__ spins, int L0, int L1, int code, int c, uint64_t sweep_add,
    const uint32_t* __restrict__ tabs, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ sids, const uint64_t* __restrict__ sweeps, int64_t n_meas,
    int64_t meas, const int64_t* __restrict__ n_meas_rep, int32_t* __restrict__ de_out,
    int32_t* __restrict__ dm_out, int32_t* __restrict__ acc_out)
{
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x);
    const uint32_t bx = lb - (uint32_t)rep * gridDim.x;
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;  // whole workgroup
    const int G = L0 >> 6;
    const int W = L1 * G;
    const int idx = (int)(bx * blockDim.x + threadIdx.x);
    Delta d{0, 0, 0};
    if (idx < W) {
        const int y = (int)fdiv((uint32_t)idx, make_fastdiv((uint32_t)G)), w = idx - y * G;
        const int p = (y + c) & 1;
        const int yp = y + 1 == L1 ? 0 : y + 1;
        const int ym = y == 0 ? L1 - 1 : y - 1;
        uint32_t own, same, sh, up, dn, other;
        int8_t* row8 = nullptr;
        uint32_t* own_word = nullptr;
        if (LAYOUT == 1) {
            uint32_t* S = (uint32_t*)spins + (int64_t)rep * 2 * W;
            const uint32_t* So = S + (1 - c) * W;
            own_word = S + c * W + idx;
            own = *own_word;
            same = So[idx];
            if (p == 0) sh = (same << 1) | (So[y * G + (w == 0 ? G - 1 : w - 1)] >> 31);
            else sh = (same >> 1) | (So[y * G + (w == G - 1 ? 0 : w + 1)] << 31);
            up = So[yp * G + w];
            dn = So[ym * G + w];
            other = 0;
        } else {
            int8_t* S = (int8_t*)spins + (int64_t)rep * L0 * L1;
            row8 = S + (int64_t)y * L0 + 64 * w;
            uint32_t ev, od;
            load_segment(row8, ev, od);
            own = p == 0 ? ev : od;
            same = p == 0 ? od : ev;
            other = same;
            if (p == 0) {
                const int8_t b = S[(int64_t)y * L0 + (w == 0 ? L0 - 1 : 64 * w - 1)];
                sh = (same << 1) | (b > 0 ? 1u : 0u);
            } else {
                const int8_t b = S[(int64_t)y * L0 + (w == G - 1 ? 0 : 64 * w + 64)];
                sh = (same >> 1) | (b > 0 ? 0x80000000u : 0u);
            }
            load_segment(S + (int64_t)yp * L0 + 64 * w, ev, od);
            up = p == 0 ? ev : od;
            load_segment(S + (int64_t)ym * L0 + 64 * w, ev, od);
            dn = p == 0 ? ev : od;
        }
        const uint32_t nb[4] = {same, sh, up, dn};
        const uint64_t sweep = sweeps[rep] + sweep_add;
        const uint32_t nw = update_word<4, NC>(code, own, nb, ~0u, tabs + (int64_t)TAB_WORDS * rep,
                                               keys[2 * rep], keys[2 * rep + 1], sids[rep], sweep,
                                               (uint32_t)c, (uint32_t)idx, d);
        if (nw != own) {
            if (LAYOUT == 1) {
                *own_word = nw;
            } else {
                const uint32_t ev = p == 0 ? nw : other, od = p == 0 ? other : nw;
                uint4* v = (uint4*)row8;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint4 q;
                    q.x = pack_pair_u32(ev, od, 8 * k + 0);
                    q.y = pack_pair_u32(ev, od, 8 * k + 2);
                    q.z = pack_pair_u32(ev, od, 8 * k + 4);
                    q.w = pack_pair_u32(ev, od, 8 * k + 6);
                    v[k] = q;
                }
            }
        }
    }
    wide_deltas(d, de_out, dm_out, acc_out, (int64_t)rep * n_meas + meas);
}

// Wide form on cubic lattices with L0 % 64 == 0 (int8): as ising2d_half_wide,
// a lane takes the 32 colour-c sites of one 64-site row segment, read as four
// 16-byte words, plus the segments of the rows at +-y and +-z and the one
// x-neighbour outside its segment.  Word idx = r G + w (row r = y + L1 z) is
// the plain form's word index, so the uniforms, and the chains, are the same;
// the neighbour order differs, which the symmetric count does not see.
template <int NC>
__global__ __launch_bounds__(WIDE_LANES) void ising3d_half_wide(
    int8_t* __restrict__ spins, int L0, int L1, int L2, int code, int c, uint64_t sweep_add,
    const uint32_t* __restrict__ tabs, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ sids, const uint64_t* __restrict__ sweeps, int64_t n_meas,
    int64_t meas, const int64_t* __restrict__ n_meas_rep, int32_t* __restrict__ de_out,
    int32_t* __restrict__ dm_out, int32_t* __restrict__ acc_out)
{
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x);
    const uint32_t bx = lb - (uint32_t)rep * gridDim.x;
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;  // whole workgroup
    const int G = L0 >> 6;
    const int rows = L1 * L2;
    const int W = rows * G;
    const int idx = (int)(bx * blockDim.x + threadIdx.x);
    Delta d{0, 0, 0};
    if (idx < W) {
        const int r = (int)fdiv((uint32_t)idx, make_fastdiv((uint32_t)G)), w = idx - r * G;
        const int z = (int)fdiv((uint32_t)r, make_fastdiv((uint32_t)L1)), y = r - z * L1;
        const int p = (y + z + c) & 1;
        const int yp = y + 1 == L1 ? 0 : y + 1, ym = y == 0 ? L1 - 1 : y - 1;
        const int zp = z + 1 == L2 ? 0 : z + 1, zm = z == 0 ? L2 - 1 : z - 1;
        int8_t* S = spins + (int64_t)rep * L0 * rows;
        int8_t* row8 = S + (int64_t)r * L0 + 64 * w;
        uint32_t ev, od;
        load_segment(row8, ev, od);
        const uint32_t own = p == 0 ? ev : od;
        const uint32_t same = p == 0 ? od : ev;
        uint32_t sh;
        if (p == 0) {
            const int8_t b = S[(int64_t)r * L0 + (w == 0 ? L0 - 1 : 64 * w - 1)];
            sh = (same << 1) | (b > 0 ? 1u : 0u);
        } else {
            const int8_t b = S[(int64_t)r * L0 + (w == G - 1 ? 0 : 64 * w + 64)];
            sh = (same >> 1) | (b > 0 ? 0x80000000u : 0u);
        }
        uint32_t nb[6] = {same, sh, 0u, 0u, 0u, 0u};
        const int nrow[4] = {z * L1 + yp, z * L1 + ym, zp * L1 + y, zm * L1 + y};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            load_segment(S + (int64_t)nrow[k] * L0 + 64 * w, ev, od);
            nb[2 + k] = p == 0 ? ev : od;
        }
        const uint64_t sweep = sweeps[rep] + sweep_add;
        const uint32_t nw = update_word<6, NC>(code, own, nb, ~0u, tabs + (int64_t)TAB_WORDS * rep,
                                               keys[2 * rep], keys[2 * rep + 1], sids[rep], sweep,
                                               (uint32_t)c, (uint32_t)idx, d);
        if (nw != own) {
            const uint32_t nev = p == 0 ? nw : same, nod = p == 0 ? same : nw;
            uint4* v = (uint4*)row8;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint4 q;
                q.x = pack_pair_u32(nev, nod, 8 * k + 0);
                q.y = pack_pair_u32(nev, nod, 8 * k + 2);
                q.z = pack_pair_u32(nev, nod, 8 * k + 4);
                q.w = pack_pair_u32(nev, nod, 8 * k + 6);
                v[k] = q;
            }
        }
    }
    wide_deltas(d, de_out, dm_out, acc_out, (int64_t)rep * n_meas + meas);
}

// ---------------------------------------------------------------------------
// Initial states, exact E_b / M, layout conversion.

__global__ void ising_init_kernel(int8_t* __restrict__ s, int64_t N, int R,
                                  const uint32_t* __restrict__ keys,
                                  const uint32_t* __restrict__ sids, int start)
{
    const int rep = blockIdx.y;
    if (rep >= R) return;
    const uint32_t k0 = keys[2 * rep], k1 = keys[2 * rep + 1], sid = sids[rep];
    for (int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; blk * 4 < N;
         blk += (int64_t)gridDim.x * blockDim.x) {
        uint32_t w[4] = {0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u};
        if (start == 0) {
            const u32x4 r = draw(k0, k1, sid, 0, 0, TAG_INIT, (uint32_t)blk);
            w[0] = r.x; w[1] = r.y; w[2] = r.z; w[3] = r.w;
        } else if (start == 2) {
            w[0] = w[1] = w[2] = w[3] = 0;
        }
        for (int k = 0; k < 4 && blk * 4 + k < N; ++k)
            s[(int64_t)rep * N + blk * 4 + k] = (w[k] >> 31) ? (int8_t)1 : (int8_t)-1;
    }
}

// E_b and M of each replica.  Workgroup b takes `chunk` sites of replica
// b / nchunk and adds its integer sums atomically onto sums zeroed by
// ising_measure_zero, so they do not depend on the order.  IX is uint32_t for
// replicas of fewer than 2^31 sites, whose coordinates then cost 32-bit
// divisions.
__global__ void ising_measure_zero(int32_t* e_run, int32_t* m_run, int n_rep)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_rep) {
        e_run[t] = 0;
        m_run[t] = 0;
    }
}

template <typename IX>
__global__ __launch_bounds__(256) void ising_measure_kernel(const int8_t* __restrict__ s, Geo g,
                                                            int64_t N, int64_t chunk,
                                                            int64_t nchunk,
                                                            int32_t* __restrict__ e_run,
                                                            int32_t* __restrict__ m_run)
{
    __shared__ int red[2 * 4];
    const int64_t rep = blockIdx.x / nchunk, part = blockIdx.x % nchunk;
    const int8_t* sr = s + rep * N;
    const IX L0 = (IX)g.L0, L1 = (IX)g.L1, L2 = (IX)g.L2;
    const IX st1 = L0, st2 = L0 * L1;
    const IX i1 = (IX)std::min<int64_t>(N, (part + 1) * chunk);
    int e = 0, m = 0;
    for (IX i = (IX)(part * chunk) + threadIdx.x; i < i1; i += blockDim.x) {
        const int si = sr[i];
        m += si;
        const IX q0 = i / L0;
        const IX x0 = i - q0 * L0;
        int nbsum = sr[x0 + 1 == L0 ? i - x0 : i + 1];
        if (g.d >= 2) {
            const IX q1 = q0 / L1;
            const IX x1 = q0 - q1 * L1;
            nbsum += sr[x1 + 1 == L1 ? i - x1 * st1 : i + st1];
            if (g.d >= 3) {
                const IX x2 = q1 % L2;
                nbsum += sr[x2 + 1 == L2 ? i - x2 * st2 : i + st2];
            }
        }
        e -= si * nbsum;
    }
    e = wave_sum(e);
    m = wave_sum(m);
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = e;
        red[2 * (threadIdx.x >> 6) + 1] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int se = 0, sm = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            se += red[2 * k];
            sm += red[2 * k + 1];
        }
        if (se) atomicAdd(e_run + rep, se);
        if (sm) atomicAdd(m_run + rep, sm);
    }
}

// int8 [R][N] <-> packed [R][2][rows][G]; direction 0 packs, 1 unpacks.
__global__ void ising_convert_kernel(int8_t* __restrict__ s8, uint32_t* __restrict__ sp, Geo g,
                                     int R, int direction)
{
    const int64_t W = g.rows * g.G;
    const int64_t total = (int64_t)R * 2 * W;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t rep = t / (2 * W);
        const int64_t rem = t - rep * 2 * W;
        const int c = (int)(rem / W);
        const int64_t gidx = rem - (int64_t)c * W;
        const int64_t r = gidx / g.G;
        const int w = (int)(gidx - r * g.G);
        int x1 = 0, x2 = 0;
        if (g.d >= 2) {
            x1 = (int)(r % g.L1);
            x2 = (int)(r / g.L1);
        }
        const int p = (x1 + x2 + c) & 1;
        const int nbits = min(32, g.M - 32 * w);
        int8_t* row = s8 + rep * g.rows * g.L0 + r * g.L0;
        if (direction == 0) {
            uint32_t v = 0;
            for (int b = 0; b < nbits; ++b) v |= (uint32_t)(row[2 * (32 * w + b) + p] > 0) << b;
            sp[t] = v;
        } else {
            const uint32_t v = sp[t];
            for (int b = 0; b < nbits; ++b)
                row[2 * (32 * w + b) + p] = ((v >> b) & 1u) ? (int8_t)1 : (int8_t)-1;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.

static uint64_t threshold(double p)
{
    if (!(p < 1.0)) return 1ull << 32;
    if (!(p > 0.0)) return 0;
    return (uint64_t)std::ceil(p * 4294967296.0);
}

// critmc/kernels/ising.py:47-65 turned into class tables (see bitserial_below).
static void build_tables(double beta, double h, int z, uint32_t* tab, int* nc_met, int* nc_hb)
{
    std::fill(tab, tab + TAB_WORDS, 0u);
    uint32_t* met = tab;
    uint32_t* hb = tab + 32;
    int n = 0;
    for (int nal = 0; nal <= z; ++nal) {
        const int a = 2 * nal - z;
        uint64_t T[2];
        for (int t = 0; t < 2; ++t) {
            const int spin = 2 * t - 1;
            const double d_e = 2.0 * a + 2.0 * h * spin;
            const double p = d_e <= 0.0 ? 1.0 : std::exp(-beta * d_e);
            T[t] = threshold(p);
        }
        if (T[0] == T[1]) {
            if (T[0] < (1ull << 32)) {
                met[1 + n] = (uint32_t)nal;
                met[TAB_T + n] = (uint32_t)T[0];
                ++n;
            }
        } else {
            for (int t = 0; t < 2; ++t) {
                if (T[t] < (1ull << 32)) {
                    met[1 + n] = (uint32_t)nal | ((uint32_t)(t == 0 ? 1 : 2) << 8);
                    met[TAB_T + n] = (uint32_t)T[t];
                    ++n;
                }
            }
        }
    }
    met[0] = (uint32_t)n;
    *nc_met = n;
    n = 0;
    for (int nup = 0; nup <= z; ++nup) {
        const int f = 2 * nup - z;
        const uint64_t T = threshold(1.0 / (1.0 + std::exp(-2.0 * beta * (f + h))));
        if (T < (1ull << 32)) {
            hb[1 + n] = (uint32_t)nup;
            hb[TAB_T + n] = (uint32_t)T;
            ++n;
        }
    }
    hb[0] = (uint32_t)n;
    *nc_hb = n;
}

static int geo_from(int ndim, const int32_t* dims, Geo* g)
{
    CM_REQUIRE(ndim >= 1 && ndim <= 3, "ndim must be 1, 2 or 3 (got %d)", ndim);
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) {
        CM_REQUIRE(dims[a] >= 2 && dims[a] % 2 == 0,
                   "checkerboard sweeps need even extents >= 2 (axis %d has %d)", a, dims[a]);
        N *= dims[a];
    }
    CM_REQUIRE(N <= (int64_t)1 << 30, "lattice too large (%lld sites)", (long long)N);
    g->d = ndim;
    g->L0 = dims[0];
    g->L1 = ndim >= 2 ? dims[1] : 1;
    g->L2 = ndim >= 3 ? dims[2] : 1;
    g->M = g->L0 / 2;
    g->G = (g->M + 31) / 32;
    g->rows = N / g->L0;
    return CM_OK;
}

static int pick_nc(int nc)
{
    if (nc <= 2) return 2;
    if (nc <= 4) return 4;
    if (nc <= 7) return 7;
    return NC_LIMIT;
}

template <int Z, int NC>
static void launch_plain(int8_t* s, const Geo& g, int code, int colour, uint64_t sweep,
                         uint32_t k0, uint32_t k1, uint32_t sid, const uint32_t* tabs,
                         int32_t* de, int32_t* dm, int32_t* ac, hipStream_t st)
{
    const int64_t nthr = g.rows * g.G;
    const int blocks = (int)((nthr + 255) / 256);
    hipLaunchKernelGGL((ising_plain_half<Z, NC>), dim3(blocks), dim3(256), 0, st, s, g, code,
                       colour, sweep, k0, k1, sid, tabs, de, dm, ac);
}

template <int Z>
static void launch_plain_nc(int nc, int8_t* s, const Geo& g, int code, int colour,
                            uint64_t sweep, uint32_t k0, uint32_t k1, uint32_t sid,
                            const uint32_t* tabs, int32_t* de, int32_t* dm, int32_t* ac,
                            hipStream_t st)
{
    switch (nc) {
        case 2: launch_plain<Z, 2>(s, g, code, colour, sweep, k0, k1, sid, tabs, de, dm, ac, st); break;
        case 4: launch_plain<Z, 4>(s, g, code, colour, sweep, k0, k1, sid, tabs, de, dm, ac, st); break;
        case 7: launch_plain<Z, 7>(s, g, code, colour, sweep, k0, k1, sid, tabs, de, dm, ac, st); break;
        default: launch_plain<Z, NC_LIMIT>(s, g, code, colour, sweep, k0, k1, sid, tabs, de, dm, ac, st); break;
    }
}

template <int Z, int NC>
static void launch_any(const dim3& grid, hipStream_t st, int8_t* s, const Geo& g, int code,
                       int colour, uint64_t sw, const uint32_t* tabs, const uint32_t* keys,
                       const uint32_t* sids, const uint64_t* sweeps, int64_t n_meas, int64_t meas,
                       const int64_t* n_meas_rep, int32_t* e_out, int32_t* m_out, int32_t* acc_out)
{
    hipLaunchKernelGGL((ising_any_half<Z, NC>), grid, dim3(256), 0, st, s, g, code, colour, sw,
                       tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out);
}

template <int NC>
static void launch_3d(const dim3& grid, hipStream_t st, int8_t* s, const Geo& g, int code,
                      int colour, uint64_t sw, const uint32_t* tabs, const uint32_t* keys,
                      const uint32_t* sids, const uint64_t* sweeps, int64_t n_meas, int64_t meas,
                      const int64_t* n_meas_rep, int32_t* e_out, int32_t* m_out, int32_t* acc_out)
{
    hipLaunchKernelGGL((ising3d_half_wide<NC>), grid, dim3(WIDE_LANES), 0, st, s, g.L0, g.L1, g.L2, code,
                       colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out,
                       m_out, acc_out);
}

static void launch_3d_nc(int nc, const dim3& grid, hipStream_t st, int8_t* s, const Geo& g,
                         int code, int colour, uint64_t sw, const uint32_t* tabs,
                         const uint32_t* keys, const uint32_t* sids, const uint64_t* sweeps,
                         int64_t n_meas, int64_t meas, const int64_t* n_meas_rep, int32_t* e_out,
                         int32_t* m_out, int32_t* acc_out)
{
    switch (nc) {
        case 2: launch_3d<2>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        case 4: launch_3d<4>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        case 7: launch_3d<7>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        default: launch_3d<NC_LIMIT>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
    }
}

template <int Z>
static void launch_any_nc(int nc, const dim3& grid, hipStream_t st, int8_t* s, const Geo& g,
                          int code, int colour, uint64_t sw, const uint32_t* tabs,
                          const uint32_t* keys, const uint32_t* sids, const uint64_t* sweeps,
                          int64_t n_meas, int64_t meas, const int64_t* n_meas_rep, int32_t* e_out,
                          int32_t* m_out, int32_t* acc_out)
{
    switch (nc) {
        case 2: launch_any<Z, 2>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        case 4: launch_any<Z, 4>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        case 7: launch_any<Z, 7>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        default: launch_any<Z, NC_LIMIT>(grid, st, s, g, code, colour, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
    }
}

template <int LAYOUT, int NC, bool ZF = false>
static hipError_t launch_batched(void* spins, int R, const Geo& g, const uint32_t* tabs,
                                 const uint32_t* keys, const uint32_t* sids, uint64_t* sweeps,
                                 const Sched& sc, int64_t n_meas, int32_t* e_run, int32_t* m_run,
                                 int32_t* e_out, int32_t* m_out, int32_t* acc_out,
                                 const int64_t* n_meas_rep, int block, int grid, hipStream_t st)
{
    const size_t lds = (size_t)2 * g.L1 * g.G * 4 + 3 * 16 * 4 + (3 * ISING_QCAP + 2) * 4;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)ising2d_batched<LAYOUT, NC, ZF>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((ising2d_batched<LAYOUT, NC, ZF>), dim3(grid), dim3(block), lds, st, spins, R,
                       g.L0, g.L1, tabs, keys, sids, sweeps, sc, n_meas, e_run, m_run, e_out,
                       m_out, acc_out, n_meas_rep);
    return hipGetLastError();
}

template <int LAYOUT>
static hipError_t launch_batched_nc(int nc, void* spins, int R, const Geo& g,
                                    const uint32_t* tabs, const uint32_t* keys,
                                    const uint32_t* sids, uint64_t* sweeps, const Sched& sc,
                                    int64_t n_meas, int32_t* e_run, int32_t* m_run,
                                    int32_t* e_out, int32_t* m_out, int32_t* acc_out,
                                    const int64_t* n_meas_rep, int block, int grid, hipStream_t st,
                                    bool zf)
{
    if (nc == 2 && zf)
        return launch_batched<LAYOUT, 2, true>(spins, R, g, tabs, keys, sids, sweeps, sc, n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block, grid, st);
    switch (nc) {
        case 2: return launch_batched<LAYOUT, 2>(spins, R, g, tabs, keys, sids, sweeps, sc, n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block, grid, st);
        case 4: return launch_batched<LAYOUT, 4>(spins, R, g, tabs, keys, sids, sweeps, sc, n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block, grid, st);
        case 7: return launch_batched<LAYOUT, 7>(spins, R, g, tabs, keys, sids, sweeps, sc, n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block, grid, st);
        default: return launch_batched<LAYOUT, NC_LIMIT>(spins, R, g, tabs, keys, sids, sweeps, sc, n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block, grid, st);
    }
}

static bool batched_ok(const Geo& g)
{
    return g.d == 2 && (g.M <= 32 || g.M % 32 == 0) && (size_t)2 * g.L1 * g.G * 4 <= 96 * 1024;
}

static bool wide_ok(const Geo& g) { return g.d == 2 && g.L0 % 64 == 0; }

// Auto choice between the batched and wide forms, from measured crossovers on
// MI355X: the batched form runs one workgroup per replica, so below a few
// hundred replicas of a large lattice most CUs would idle; the wide form
// re-reads neighbour rows from HBM every half sweep, which costs more in the
// int8 layout (8 B per site) than packed (0.6 B per site).
static bool prefer_wide(const Geo& g, int layout, int n_rep)
{
    const int64_t N = g.rows * g.L0;
    if (!(g.d == 2 && g.L0 % 64 == 0) || N < 65536) return false;
    if (layout == 1) return n_rep < 512;
    return n_rep < 64 || (N >= 262144 && n_rep < 256);
}

template <int LAYOUT, int NC>
static void launch_wide(const dim3& grid, hipStream_t st, void* spins, const Geo& g, int code,
                        int c, uint64_t sw, const uint32_t* tabs, const uint32_t* keys,
                        const uint32_t* sids, const uint64_t* sweeps, int64_t n_meas, int64_t meas,
                        const int64_t* n_meas_rep, int32_t* e_out, int32_t* m_out, int32_t* acc_out)
{
    hipLaunchKernelGGL((ising2d_half_wide<LAYOUT, NC>), grid, dim3(WIDE_LANES), 0, st, spins, g.L0, g.L1,
                       code, c, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out,
                       m_out, acc_out);
}

template <int LAYOUT>
static void launch_wide_nc(int nc, const dim3& grid, hipStream_t st, void* spins, const Geo& g,
                           int code, int c, uint64_t sw, const uint32_t* tabs, const uint32_t* keys,
                           const uint32_t* sids, const uint64_t* sweeps, int64_t n_meas,
                           int64_t meas, const int64_t* n_meas_rep, int32_t* e_out, int32_t* m_out,
                           int32_t* acc_out)
{
    switch (nc) {
        case 2: launch_wide<LAYOUT, 2>(grid, st, spins, g, code, c, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        case 4: launch_wide<LAYOUT, 4>(grid, st, spins, g, code, c, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        case 7: launch_wide<LAYOUT, 7>(grid, st, spins, g, code, c, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
        default: launch_wide<LAYOUT, NC_LIMIT>(grid, st, spins, g, code, c, sw, tabs, keys, sids, sweeps, n_meas, meas, n_meas_rep, e_out, m_out, acc_out); break;
    }
}

// Schedule rows (code, count) of a checkerboard run, validated.
static int parse_sched(const int64_t* schedule, int n_rows, Sched* sc)
{
    CM_REQUIRE(n_rows >= 1 && n_rows <= 8, "schedule must have 1..8 rows");
    sc->n = n_rows;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r], count = schedule[2 * r + 1];
        CM_REQUIRE(code == CODE_METROPOLIS || code == CODE_HEATBATH,
                   "update code %lld has no checkerboard Ising kernel", (long long)code);
        CM_REQUIRE(count >= 1 && count <= (1 << 20), "repetition count out of range");
        sc->code[r] = (int32_t)code;
        sc->count[r] = (int32_t)count;
    }
    return CM_OK;
}

// Class tables of every chain (tabs: n_rep x TAB_WORDS) and the record the
// launch needs: info[0] = the largest class count a row of this schedule
// reads, info[1] = 1 when every row is Metropolis and every table is the
// zero-field z = 4 one (the kernel compiled for that rule alone applies).
static int prepare_tables(int z, int n_rep, const double* betas, double h, const Sched& sc,
                          uint32_t* tabs, int32_t* info)
{
    bool uses_met = false, uses_hb = false;
    for (int r = 0; r < sc.n; ++r) {
        uses_met |= sc.code[r] == CODE_METROPOLIS;
        uses_hb |= sc.code[r] == CODE_HEATBATH;
    }
    int nc_max = 1;
    bool met34 = z == 4 && uses_met && !uses_hb;
    for (int r = 0; r < n_rep; ++r) {
        CM_REQUIRE(std::isfinite(betas[r]), "beta must be finite");
        uint32_t* tab = tabs + (size_t)r * TAB_WORDS;
        int a, b;
        build_tables(betas[r], h, z, tab, &a, &b);
        if (uses_met) nc_max = std::max(nc_max, a);
        if (uses_hb) nc_max = std::max(nc_max, b);
        met34 = met34 && table_is_met34(tab);
    }
    info[0] = nc_max;
    info[1] = met34 ? 1 : 0;
    return CM_OK;
}

// Whether cm_ising_run with form 0 on int8 spins only launches kernels, so
// that its launches can be captured: form 0 there is always the batched or a
// wide form, never the plain form (which reads the host sweep counters and
// synchronises the stream), so every valid lattice qualifies.
bool ising_rows_capturable(int32_t ndim, const int32_t* dims, int32_t n_rep)
{
    (void)n_rep;  // on int8 spins form 0 is always the batched or a wide form
    Geo g;
    return geo_from(ndim, dims, &g) == CM_OK;
}

}  // namespace cm

using namespace cm;

extern "C" {

int64_t cm_ising_workspace_bytes(int32_t n_rep) { return (int64_t)n_rep * TAB_WORDS * 4; }

int cm_ising_tables(double beta, double h, int32_t z, uint32_t* tab_out)
{
    CM_REQUIRE(z == 2 || z == 4 || z == 6, "z must be 2, 4 or 6");
    int a, b;
    build_tables(beta, h, z, tab_out, &a, &b);
    return CM_OK;
}

int cm_ising_init(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                  const uint32_t* keys, const uint32_t* stream_ids, int32_t start, void* stream)
{
    Geo g;
    if (int rc = geo_from(ndim, dims, &g)) return rc;
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    CM_REQUIRE(start >= 0 && start <= 2, "start must be 0 (random), 1 (up) or 2 (down)");
    const int64_t N = g.rows * g.L0;
    const int64_t blocks_needed = (N / 4 + 255) / 256;
    dim3 grid((unsigned)std::min<int64_t>(std::max<int64_t>(blocks_needed, 1), 1024), n_rep);
    hipLaunchKernelGGL(ising_init_kernel, grid, dim3(256), 0, (hipStream_t)stream, (int8_t*)spins,
                       N, n_rep, keys, stream_ids, start);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_ising_measure(const void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                     int32_t* e_run, int32_t* m_run, void* stream)
{
    Geo g;
    if (int rc = geo_from(ndim, dims, &g)) return rc;
    CM_REQUIRE(n_rep >= 1, "n_rep must be positive");
    const int64_t N = g.rows * g.L0;
    int64_t chunk = 4096;
    while ((int64_t)n_rep * ((N + chunk - 1) / chunk) > (int64_t)1 << 22) chunk *= 2;
    const int64_t nchunk = (N + chunk - 1) / chunk;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(ising_measure_zero, dim3((n_rep + 255) / 256), dim3(256), 0, st, e_run,
                       m_run, n_rep);
    if (N < ((int64_t)1 << 31))
        hipLaunchKernelGGL(ising_measure_kernel<uint32_t>, dim3((unsigned)(n_rep * nchunk)),
                           dim3(256), 0, st, (const int8_t*)spins, g, N, chunk, nchunk, e_run,
                           m_run);
    else
        hipLaunchKernelGGL(ising_measure_kernel<uint64_t>, dim3((unsigned)(n_rep * nchunk)),
                           dim3(256), 0, st, (const int8_t*)spins, g, N, chunk, nchunk, e_run,
                           m_run);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_ising_convert(void* spins_i8, void* spins_packed, int32_t ndim, const int32_t* dims,
                     int32_t n_rep, int32_t direction, void* stream)
{
    Geo g;
    if (int rc = geo_from(ndim, dims, &g)) return rc;
    CM_REQUIRE(direction == 0 || direction == 1, "direction must be 0 (pack) or 1 (unpack)");
    const int64_t total = (int64_t)n_rep * 2 * g.rows * g.G;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(ising_convert_kernel, dim3(std::max(blocks, 1)), dim3(256), 0,
                       (hipStream_t)stream, (int8_t*)spins_i8, (uint32_t*)spins_packed, g, n_rep,
                       direction);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_ising_prepare(int32_t ndim, const int32_t* dims, int32_t n_rep, const double* betas,
                     double h, const int64_t* schedule, int32_t n_rows, void* workspace,
                     int64_t workspace_bytes, int32_t* info, void* stream)
{
    Geo g;
    if (int rc = geo_from(ndim, dims, &g)) return rc;
    CM_REQUIRE(n_rep >= 1, "n_rep must be positive");
    CM_REQUIRE(workspace_bytes >= cm_ising_workspace_bytes(n_rep), "workspace too small");
    CM_REQUIRE(std::isfinite(h), "field h must be finite");
    Sched sc{};
    if (int rc = parse_sched(schedule, n_rows, &sc)) return rc;
    std::vector<uint32_t> tabs((size_t)n_rep * TAB_WORDS);
    if (int rc = prepare_tables(2 * g.d, n_rep, betas, h, sc, tabs.data(), info)) return rc;
    CM_HIP_CHECK(upload_async(workspace, tabs.data(), tabs.size() * 4, (hipStream_t)stream));
    return CM_OK;
}

int cm_ising_run(void* spins, int32_t layout, int32_t ndim, const int32_t* dims, int32_t n_rep,
                 const double* betas, double h, const uint32_t* keys, const uint32_t* stream_ids,
                 uint64_t* sweeps, const uint64_t* sweeps_host, const int64_t* schedule,
                 int32_t n_rows, int64_t n_meas, int32_t* e_run, int32_t* m_run, int32_t* e_out,
                 int32_t* m_out, int32_t* acc_out, void* workspace, int64_t workspace_bytes,
                 int32_t form, int32_t block, int32_t grid, const int64_t* n_meas_rep,
                 const int32_t* prepared, void* stream)
{
    Geo g;
    if (int rc = geo_from(ndim, dims, &g)) return rc;
    CM_REQUIRE(n_rep >= 1, "n_rep must be positive");
    CM_REQUIRE(layout == 0 || layout == 1, "layout must be 0 (int8) or 1 (packed)");
    CM_REQUIRE(n_meas >= 0, "n_meas must not be negative");
    CM_REQUIRE(workspace_bytes >= cm_ising_workspace_bytes(n_rep), "workspace too small");
    CM_REQUIRE(std::isfinite(h), "field h must be finite");
    Sched sc{};
    if (int rc = parse_sched(schedule, n_rows, &sc)) return rc;
    if (n_meas == 0) return CM_OK;
    const int z = 2 * g.d;
    hipStream_t st = (hipStream_t)stream;
    uint32_t* dtabs = (uint32_t*)workspace;
    int32_t info[2];
    if (prepared != nullptr) {  // tables already in the workspace (cm_ising_prepare)
        CM_REQUIRE(prepared[0] >= 1 && prepared[0] <= NC_LIMIT && (prepared[1] & ~1) == 0,
                   "prepared table record is not one cm_ising_prepare wrote");
        info[0] = prepared[0];
        info[1] = prepared[1];
    } else {
        std::vector<uint32_t> tabs((size_t)n_rep * TAB_WORDS);
        if (int rc = prepare_tables(z, n_rep, betas, h, sc, tabs.data(), info)) return rc;
        CM_HIP_CHECK(upload_async(dtabs, tabs.data(), tabs.size() * 4, st));
    }
    const int nc = pick_nc(info[0]);
    const bool zf_tables = info[1] != 0;

    if (form == 0) {
        // int8 lattices that neither 2D form takes (chains, cubic, 2D rows
        // that are not whole 64-site segments) go wide on the plain form's
        // words: every replica in each launch, no host synchronisation
        if (prefer_wide(g, layout, n_rep)) form = 3;
        else if (batched_ok(g)) form = 2;
        else form = layout == 0 ? 3 : 1;
    }
    if (form == 3 && !wide_ok(g)) {
        CM_REQUIRE(layout == 0, "the wide form off 2D lattices with L0 a multiple of 64 works "
                                "on int8 spins");
        CM_REQUIRE(n_rep <= 65535, "n_rep out of range");
        const int64_t total = (int64_t)n_rep * n_meas;
        const int fb = (int)std::min<int64_t>((total + 255) / 256, 4096);
        hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, e_out, total, 0);
        hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, m_out, total, 0);
        hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, acc_out, total, 0);
        const int64_t words = g.rows * g.G;
        // cubic lattices of 64-site row segments: whole segments per lane
        const bool seg3 = g.d == 3 && g.L0 % 64 == 0;
        const int wpb = seg3 ? WIDE_LANES : 256;  // lanes per workgroup
        CM_REQUIRE((words + wpb - 1) / wpb * n_rep < (1ll << 32),
                   "too many words x replicas per launch");
        const dim3 agrid((unsigned)((words + wpb - 1) / wpb), (unsigned)n_rep);
        int8_t* s = (int8_t*)spins;
        uint64_t sw = 0;
        int64_t per_meas = 0;
        for (int r = 0; r < sc.n; ++r) per_meas += sc.count[r];
        for (int64_t meas = 0; meas < n_meas; ++meas)
            for (int r = 0; r < sc.n; ++r)
                for (int rp = 0; rp < sc.count[r]; ++rp) {
                    for (int c = 0; c < 2; ++c) {
                        if (seg3)
                            launch_3d_nc(nc, agrid, st, s, g, sc.code[r], c, sw, dtabs, keys,
                                         stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out,
                                         m_out, acc_out);
                        else if (z == 2)
                            launch_any_nc<2>(nc, agrid, st, s, g, sc.code[r], c, sw, dtabs, keys,
                                             stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out,
                                             m_out, acc_out);
                        else if (z == 4)
                            launch_any_nc<4>(nc, agrid, st, s, g, sc.code[r], c, sw, dtabs, keys,
                                             stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out,
                                             m_out, acc_out);
                        else
                            launch_any_nc<6>(nc, agrid, st, s, g, sc.code[r], c, sw, dtabs, keys,
                                             stream_ids, sweeps, n_meas, meas, n_meas_rep, e_out,
                                             m_out, acc_out);
                    }
                    ++sw;
                }
        hipLaunchKernelGGL(prefix_series, dim3((n_rep + 255) / 256), dim3(256), 0, st, e_out,
                           m_out, n_meas, e_run, m_run, n_rep, n_meas_rep, sweeps,
                           (uint64_t)per_meas);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    }
    if (form == 3) {
        CM_REQUIRE(n_rep <= 65535, "n_rep out of range");
        const int64_t total = (int64_t)n_rep * n_meas;
        const int fb = (int)std::min<int64_t>((total + 255) / 256, 4096);
        hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, e_out, total, 0);
        hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, m_out, total, 0);
        hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, acc_out, total, 0);
        const int64_t words = (int64_t)g.L1 * (g.L0 / 64);
        const dim3 wgrid((unsigned)((words + WIDE_LANES - 1) / WIDE_LANES), (unsigned)n_rep);
        uint64_t sw = 0;
        int64_t per_meas = 0;
        for (int r = 0; r < sc.n; ++r) per_meas += sc.count[r];
        for (int64_t meas = 0; meas < n_meas; ++meas)
            for (int r = 0; r < sc.n; ++r)
                for (int rp = 0; rp < sc.count[r]; ++rp) {
                    for (int c = 0; c < 2; ++c) {
                        if (layout == 0)
                            launch_wide_nc<0>(nc, wgrid, st, spins, g, sc.code[r], c, sw, dtabs,
                                              keys, stream_ids, sweeps, n_meas, meas, n_meas_rep,
                                              e_out, m_out, acc_out);
                        else
                            launch_wide_nc<1>(nc, wgrid, st, spins, g, sc.code[r], c, sw, dtabs,
                                              keys, stream_ids, sweeps, n_meas, meas, n_meas_rep,
                                              e_out, m_out, acc_out);
                    }
                    ++sw;
                }
        hipLaunchKernelGGL(prefix_series, dim3((n_rep + 255) / 256), dim3(256), 0, st, e_out,
                           m_out, n_meas, e_run, m_run, n_rep, n_meas_rep, sweeps,
                           (uint64_t)per_meas);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    }
    if (form == 2) {
        CM_REQUIRE(batched_ok(g),
                   "batched form needs a 2D lattice with L0 <= 64 or L0 a multiple of 64");
        if (block <= 0) {  // one lane per word of a colour, in whole waves, at most 256
            const int64_t words = (int64_t)g.L1 * g.G;
            block = (int)std::min<int64_t>(256, std::max<int64_t>(64, (words + 63) / 64 * 64));
        }
        CM_REQUIRE(block % 64 == 0 && block <= 256, "block must be a multiple of 64, <= 256");
        if (grid <= 0) grid = n_rep;
        grid = std::min(grid, n_rep);
        // zero-field Metropolis at z = 4 in every row and replica: the kernel
        // compiled for that rule alone
        const bool zf = g.d == 2 && zf_tables;
        hipError_t err = layout == 0
            ? launch_batched_nc<0>(nc, spins, n_rep, g, dtabs, keys, stream_ids, sweeps, sc,
                                   n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block,
                                   grid, st, zf)
            : launch_batched_nc<1>(nc, spins, n_rep, g, dtabs, keys, stream_ids, sweeps, sc,
                                   n_meas, e_run, m_run, e_out, m_out, acc_out, n_meas_rep, block, grid, st, zf);
        CM_HIP_CHECK(err);
        return CM_OK;
    }
    // plain form: one chain at a time, one launch per half sweep
    CM_REQUIRE(form == 1, "form must be 0 (auto), 1 (plain), 2 (batched) or 3 (wide)");
    CM_REQUIRE(n_meas_rep == nullptr, "per-chain measurement counts need the batched form");
    CM_REQUIRE(layout == 0, "the plain form works on int8 spins");
    CM_REQUIRE(sweeps_host != nullptr, "the plain form needs the sweep counters on the host");
    std::vector<uint32_t> hkeys((size_t)2 * n_rep), hsid(n_rep);
    CM_HIP_CHECK(hipMemcpyAsync(hkeys.data(), keys, hkeys.size() * 4, hipMemcpyDeviceToHost, st));
    CM_HIP_CHECK(hipMemcpyAsync(hsid.data(), stream_ids, hsid.size() * 4, hipMemcpyDeviceToHost, st));
    CM_HIP_CHECK(hipStreamSynchronize(st));
    const int64_t N = g.rows * g.L0;
    const int64_t total = (int64_t)n_rep * n_meas;
    const int fb = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, e_out, total, 0);
    hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, m_out, total, 0);
    hipLaunchKernelGGL(fill_i32, dim3(fb), dim3(256), 0, st, acc_out, total, 0);
    for (int r = 0; r < n_rep; ++r) {
        int8_t* s = (int8_t*)spins + (int64_t)r * N;
        const uint32_t* t = dtabs + (int64_t)r * TAB_WORDS;
        uint64_t sweep = sweeps_host[r];
        for (int64_t meas = 0; meas < n_meas; ++meas) {
            const int64_t o = (int64_t)r * n_meas + meas;
            for (int row = 0; row < sc.n; ++row) {
                for (int rp = 0; rp < sc.count[row]; ++rp) {
                    for (int c = 0; c < 2; ++c) {
                        if (z == 2)
                            launch_plain_nc<2>(nc, s, g, sc.code[row], c, sweep, hkeys[2 * r],
                                               hkeys[2 * r + 1], hsid[r], t, e_out + o, m_out + o,
                                               acc_out + o, st);
                        else if (z == 4)
                            launch_plain_nc<4>(nc, s, g, sc.code[row], c, sweep, hkeys[2 * r],
                                               hkeys[2 * r + 1], hsid[r], t, e_out + o, m_out + o,
                                               acc_out + o, st);
                        else
                            launch_plain_nc<6>(nc, s, g, sc.code[row], c, sweep, hkeys[2 * r],
                                               hkeys[2 * r + 1], hsid[r], t, e_out + o, m_out + o,
                                               acc_out + o, st);
                    }
                    ++sweep;
                }
            }
        }
    }
    CM_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(prefix_series, dim3((n_rep + 255) / 256), dim3(256), 0, st, e_out, m_out,
                       n_meas, e_run, m_run, n_rep, (const int64_t*)nullptr, (uint64_t*)nullptr,
                       (uint64_t)0);
    // advance the device-side sweep counters like the batched form does
    std::vector<uint64_t> nsw(n_rep);
    int64_t per_meas = 0;
    for (int row = 0; row < sc.n; ++row) per_meas += sc.count[row];
    for (int r = 0; r < n_rep; ++r) nsw[r] = sweeps_host[r] + (uint64_t)(per_meas * n_meas);
    CM_HIP_CHECK(hipMemcpyAsync(sweeps, nsw.data(), nsw.size() * 8, hipMemcpyHostToDevice, st));
    CM_HIP_CHECK(hipStreamSynchronize(st));
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // extern "C"
