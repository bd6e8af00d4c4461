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


}  // namespace cm
