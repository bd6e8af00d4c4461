// Sequential single-site updates in the reference's own site orders, one lane
// per chain, on any lattice given by its neighbour table nbr[N][z] (pair form,
// column 2a = +e_a, 2a+1 = -e_a; critmc/lattice.py:129-134, 173-213).
//
//   code 0  metropolis          sites 0..N-1 in turn (critmc/kernels/ising.py:68-87,
//                               kernels/potts.py:44-71 with random_order False)
//   code 1  metropolis-random   N attempts at uniformly drawn sites (the same
//                               functions with random_order True, kernels/common.py:43-51)
//   code 2  heatbath            sites 0..N-1 in turn (kernels/ising.py:90-107,
//                               kernels/potts.py:74-103)
//
// Uniforms: attempt t of a sweep uses Philox block (t, sweep_lo, sweep_hi16 |
// code << 16 | TAG_SEQ << 24, stream) with the chain's key; its words are
//   Ising  code 0: w0 acceptance; code 1: w0 site floor(w0 N / 2^32), w1 acceptance;
//          code 2: w0 (new spin +1 iff w0 < ceil(p_up 2^32))
//   Potts  code 0: w0 proposal floor(w0 (q-1) / 2^32) (+1 if >= old), w1 acceptance;
//          code 1: w0 site, w1 proposal, w2 acceptance;
//          code 2: u = w0 2^-32, new state from the FP64 sums of exp(beta n_k)
//          in the reference's order
// An acceptance "u < p" is the integer test w < ceil(p 2^32) against tables the
// host builds in FP64, so the chains are bit-identical to the CPU replay.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "philox.hpp"

namespace cm {

constexpr uint32_t TAG_SEQ = 11;
constexpr int SEQ_ZMAX = 8;            // coordination numbers up to 8 (4D hypercubic)
constexpr int SEQ_TAB = 64;            // u64 words of per-chain tables
constexpr int SEQ_HB = 2 * (2 * SEQ_ZMAX + 1);  // Ising heat-bath thresholds start here
constexpr int SEQ_W = 32;              // Potts heat-bath weights (FP64) start here

struct SeqArgs {
    int8_t* spins;           // [R][N]
    const int32_t* nbr;      // [N][z]
    uint32_t N;
    int z, R, q;             // q = 0: Ising +-1, else Potts 0..q-1
    const uint64_t* tab;     // [R][SEQ_TAB]
    const uint32_t* keys;    // [R][2]
    const uint32_t* sids;    // [R]
    uint64_t* sweeps;        // [R], advanced by the kernel
    int32_t* e_run;          // [R]: E_b (Ising) or E (Potts)
    int32_t* m_run;          // [R] (Ising)
    int32_t* counts;         // [R][q] (Potts)
    int64_t* acc;            // [R] accepted moves, added to (null: not kept)
};

struct SeqRows {
    int n;
    int32_t code[8], count[8];
};

__device__ __forceinline__ u32x4 seq_block(uint32_t t, uint64_t sweep, uint32_t zc, uint32_t sid,
                                           uint32_t k0, uint32_t k1)
{
    return philox4x32_10(u32x4{t, (uint32_t)sweep, zc, sid}, k0, k1);
}

// One local sweep of an Ising chain.  Returns the accepted (changed) spins.
__device__ __forceinline__ int64_t seq_ising_sweep(const SeqArgs& a, int8_t* __restrict__ s,
                                                   const uint64_t* __restrict__ tab, int code,
                                                   uint64_t sweep, uint32_t k0, uint32_t k1,
                                                   uint32_t sid, int& e, int& m)
{
    const uint32_t N = a.N;
    const int z = a.z;
    const uint32_t zc = ctr_z(sweep, (uint32_t)code, TAG_SEQ);
    int64_t accepted = 0;
    for (uint32_t t = 0; t < N; ++t) {
        const u32x4 r = seq_block(t, sweep, zc, sid, k0, k1);
        uint32_t i = t, u = r.x;
        if (code == 1) {
            i = (uint32_t)(((uint64_t)r.x * N) >> 32);
            u = r.y;
        }
        const int32_t* nb = a.nbr + (int64_t)i * z;
        int field = 0;
        for (int k = 0; k < z; ++k) field += s[nb[k]];
        const int si = s[i];
        if (code == 2) {
            const int nw = (uint64_t)u < tab[SEQ_HB + field + z] ? 1 : -1;
            if (nw != si) {
                e += 2 * si * field;
                m += 2 * nw;
                s[i] = (int8_t)nw;
                ++accepted;
            }
        } else {
            const int al = si * field;
            if ((uint64_t)u < tab[2 * (al + z) + ((si + 1) >> 1)]) {
                s[i] = (int8_t)(-si);
                e += 2 * al;
                m -= 2 * si;
                ++accepted;
            }
        }
    }
    return accepted;
}

// One local sweep of a Potts chain; `cnt` is its [q] occupation row.
__device__ __forceinline__ int64_t seq_potts_sweep(const SeqArgs& a, int8_t* __restrict__ s,
                                                   const uint64_t* __restrict__ tab,
                                                   int32_t* __restrict__ cnt, int code,
                                                   uint64_t sweep, uint32_t k0, uint32_t k1,
                                                   uint32_t sid, int& e)
{
    const uint32_t N = a.N;
    const int z = a.z, q = a.q;
    const double* w = (const double*)(tab + SEQ_W);
    const uint32_t zc = ctr_z(sweep, (uint32_t)code, TAG_SEQ);
    int64_t accepted = 0;
    for (uint32_t t = 0; t < N; ++t) {
        const u32x4 r = seq_block(t, sweep, zc, sid, k0, k1);
        uint32_t i = t, wp = r.x, wa = r.y;
        if (code == 1) {
            i = (uint32_t)(((uint64_t)r.x * N) >> 32);
            wp = r.y;
            wa = r.z;
        }
        const int32_t* nb = a.nbr + (int64_t)i * z;
        int ns[SEQ_ZMAX];
#pragma unroll
        for (int k = 0; k < SEQ_ZMAX; ++k) ns[k] = k < z ? s[nb[k]] : -1;
        const int old = s[i];
        int nw;
        int dn;  // n_old - n_new
        if (code == 2) {
            // n_k and the weights summed in state order, as the reference does
            double total = 0.0;
            for (int k = 0; k < q; ++k) {
                int nk = 0;
#pragma unroll
                for (int c = 0; c < SEQ_ZMAX; ++c) nk += ns[c] == k;
                total += w[nk];
            }
            const double rr = ((double)r.x * 0x1p-32) * total;
            nw = q - 1;
            double cum = 0.0;
            for (int k = 0; k < q; ++k) {
                int nk = 0;
#pragma unroll
                for (int c = 0; c < SEQ_ZMAX; ++c) nk += ns[c] == k;
                cum += w[nk];
                if (rr < cum) {
                    nw = k;
                    break;
                }
            }
            if (nw == old) continue;
        } else {
            nw = (int)(((uint64_t)wp * (uint32_t)(q - 1)) >> 32);
            if (nw >= old) ++nw;
        }
        int n_old = 0, n_new = 0;
#pragma unroll
        for (int c = 0; c < SEQ_ZMAX; ++c) {
            n_old += ns[c] == old;
            n_new += ns[c] == nw;
        }
        dn = n_old - n_new;
        if (code != 2 && !(dn <= 0 || (uint64_t)wa < tab[dn])) continue;
        s[i] = (int8_t)nw;
        cnt[old] -= 1;
        cnt[nw] += 1;
        e += dn;
        ++accepted;
    }
    return accepted;
}

// `rows` applied n_meas times to every chain, one lane per chain.  With
// e_out != null each measurement records E (and M or counts, and aux) as the
// schedule interpreter of the reference does; otherwise accepted moves are
// added to a.acc.  Each repetition of a row takes one sweep-counter value.
__global__ __launch_bounds__(64) void seq_run(SeqArgs a, SeqRows rows, int64_t n_meas,
                                              int32_t* e_out, int32_t* mc_out, int64_t* aux_out)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= a.R) return;
    int8_t* s = a.spins + (int64_t)rep * a.N;
    const uint64_t* tab = a.tab + (int64_t)rep * SEQ_TAB;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    int32_t* cnt = a.q ? a.counts + (int64_t)rep * a.q : nullptr;
    uint64_t sweep = a.sweeps[rep];
    int e = a.e_run[rep];
    int m = a.q ? 0 : a.m_run[rep];
    int64_t total = 0;
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t accepted = 0;
        for (int r = 0; r < rows.n; ++r)
            for (int k = 0; k < rows.count[r]; ++k, ++sweep)
                accepted += a.q ? seq_potts_sweep(a, s, tab, cnt, rows.code[r], sweep, k0, k1, sid, e)
                                : seq_ising_sweep(a, s, tab, rows.code[r], sweep, k0, k1, sid, e, m);
        total += accepted;
        if (e_out) {
            const int64_t o = (int64_t)rep * n_meas + meas;
            e_out[o] = e;
            if (a.q) {
                for (int c = 0; c < a.q; ++c) mc_out[o * a.q + c] = cnt[c];
            } else {
                mc_out[o] = m;
            }
            aux_out[4 * o + 0] = accepted;
            aux_out[4 * o + 1] = 0;
            aux_out[4 * o + 2] = 0;
            aux_out[4 * o + 3] = 0;
        }
    }
    a.sweeps[rep] = sweep;
    a.e_run[rep] = e;
    if (!a.q) a.m_run[rep] = m;
    if (a.acc) a.acc[rep] += total;
}

// Exact E (and M or counts) of every chain from the table: one workgroup per chain.
__global__ __launch_bounds__(256) void seq_measure(SeqArgs a)
{
    __shared__ long long red[2][4];
    const int rep = blockIdx.x;
    const int8_t* s = a.spins + (int64_t)rep * a.N;
    int32_t* cnt = a.q ? a.counts + (int64_t)rep * a.q : nullptr;
    if (a.q)
        for (int k = threadIdx.x; k < a.q; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    long long e = 0, m = 0;
    for (uint32_t i = threadIdx.x; i < a.N; i += blockDim.x) {
        const int si = s[i];
        const int32_t* nb = a.nbr + (int64_t)i * a.z;
        for (int ax = 0; ax < a.z / 2; ++ax) {
            const int sj = s[nb[2 * ax]];
            e -= a.q ? (long long)(si == sj) : (long long)(si * sj);
        }
        if (a.q) atomicAdd(cnt + si, 1);
        else m += si;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        m += __shfl_xor(m, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = e;
        red[1][threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long te = 0, tm = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            te += red[0][w];
            tm += red[1][w];
        }
        a.e_run[rep] = (int32_t)te;
        if (!a.q) a.m_run[rep] = (int32_t)tm;
    }
}

}  // namespace cm
