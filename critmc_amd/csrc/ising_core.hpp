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

// Bit-serial decision state of one word: lane class masks and thresholds, the
// lanes still undecided (pending) and those already known to be below their
// threshold (below).  Lanes in no class accept unconditionally (p >= 1).
template <int NCMAX>
struct WordDecision {
    uint32_t c0, c1, c2;     // bit-sliced neighbour count
    uint32_t mask[NCMAX];
    uint32_t thr[NCMAX];
    uint32_t pending, below;
};

// Count neighbours (aligned for Metropolis, up for the heat bath) and sort the
// lanes into the table's classes.  `tabs` is the replica's 64-word table:
// [0, 32) Metropolis, [32, 64) heat bath; one 32-word table holds [0] = class
// count, [1 + k] = count value | spin selector << 8 (0 any, 1 spin -1,
// 2 spin +1), [TAB_T + k] = threshold.
template <int Z, int NCMAX>
__host__ __device__ __forceinline__ void word_classify(int code, uint32_t own, const uint32_t (&nb)[Z],
                                                       uint32_t valid, const uint32_t* __restrict__ tabs,
                                                       WordDecision<NCMAX>& w)
{
    const uint32_t* tab = tabs;
    if (code == CODE_METROPOLIS) {
        uint32_t in[Z];
#pragma unroll
        for (int k = 0; k < Z; ++k) in[k] = ~(own ^ nb[k]);  // aligned neighbours
        count_bits<Z>(in, w.c0, w.c1, w.c2);
    } else {
        count_bits<Z>(nb, w.c0, w.c1, w.c2);  // neighbours up
        tab = tabs + 32;
    }
    const uint32_t ncls = tab[0];
    uint32_t pending = 0;
#pragma unroll
    for (int k = 0; k < NCMAX; ++k) {
        uint32_t mk = 0;
        if ((uint32_t)k < ncls) {
            const uint32_t d = tab[1 + k];
            const uint32_t v = d & 0xffu, sel = d >> 8;
            mk = (w.c0 ^ ((v & 1u) ? 0u : ~0u)) & (w.c1 ^ ((v & 2u) ? 0u : ~0u)) &
                 (w.c2 ^ ((v & 4u) ? 0u : ~0u));
            if (sel == 1u) mk &= ~own;
            else if (sel == 2u) mk &= own;
            mk &= valid;
        }
        w.mask[k] = mk;
        w.thr[k] = tab[TAB_T + k];
        pending |= mk;
    }
    w.pending = pending;
    w.below = valid & ~pending;
}

// One update rule's class table decoded once per half sweep: for class k,
// (c_b ^ x_b[k]) is all-ones in the lanes whose count bit b equals bit b of the
// class's count value, and (own ^ sa[k]) | sb[k] keeps the lanes of the
// class's spin (all lanes for selector 0).  Unused classes get count value 7,
// which a count of at most 6 neighbours never takes, so their masks are empty
// without a branch.
template <int NCMAX>
struct ClassTab {
    uint32_t x0[NCMAX], x1[NCMAX], x2[NCMAX], sa[NCMAX], sb[NCMAX], thr[NCMAX];
    uint32_t anysel;  // nonzero when some class selects a spin
};

template <int NCMAX>
__host__ __device__ __forceinline__ void decode_classes(const uint32_t* __restrict__ tab,
                                                        ClassTab<NCMAX>& t)
{
    const uint32_t ncls = tab[0];
    t.anysel = 0;
#pragma unroll
    for (int k = 0; k < NCMAX; ++k) {
        uint32_t v = 7u, sel = 0u, thr = 0u;
        if ((uint32_t)k < ncls) {
            const uint32_t d = tab[1 + k];
            v = d & 0xffu;
            sel = d >> 8;
            thr = tab[TAB_T + k];
        }
        t.x0[k] = (v & 1u) ? 0u : ~0u;
        t.x1[k] = (v & 2u) ? 0u : ~0u;
        t.x2[k] = (v & 4u) ? 0u : ~0u;
        t.sa[k] = sel == 2u ? 0u : ~0u;
        t.sb[k] = sel == 0u ? ~0u : 0u;
        t.thr[k] = thr;
        t.anysel |= sel;
    }
}

// word_classify with the table already decoded (same masks for valid lanes;
// a class mask may include invalid lanes, which are never pending, so no
// decision and no count sees them).  SEL: some class selects a spin
// (Metropolis in a field); without it the spin selection is skipped.
template <int Z, int NCMAX, bool SEL>
__host__ __device__ __forceinline__ void word_classify_decoded(int code, uint32_t own,
                                                               const uint32_t (&nb)[Z],
                                                               uint32_t valid,
                                                               const ClassTab<NCMAX>& t,
                                                               WordDecision<NCMAX>& w)
{
    if (code == CODE_METROPOLIS) {
        uint32_t in[Z];
#pragma unroll
        for (int k = 0; k < Z; ++k) in[k] = ~(own ^ nb[k]);  // aligned neighbours
        count_bits<Z>(in, w.c0, w.c1, w.c2);
    } else {
        count_bits<Z>(nb, w.c0, w.c1, w.c2);  // neighbours up
    }
#pragma unroll
    for (int k = 0; k < NCMAX; ++k) {
        w.mask[k] = (w.c0 ^ t.x0[k]) & (w.c1 ^ t.x1[k]) & (w.c2 ^ t.x2[k]);
        w.thr[k] = t.thr[k];
    }
    if constexpr (SEL) {
#pragma unroll
        for (int k = 0; k < NCMAX; ++k) w.mask[k] &= (own ^ t.sa[k]) | t.sb[k];
    }
    uint32_t pending = 0;
#pragma unroll
    for (int k = 0; k < NCMAX; ++k) pending |= w.mask[k];
    w.pending = pending & valid;
    w.below = valid & ~pending;
}

// Metropolis at z = 4 without a field, whose classes are exactly aligned
// counts 3 and 4 (every beta > 0): a count of at most 4 has c2 set only at 4,
// and c0 = c1 = 1 only at 3, so the class masks are c0 & c1 and c2.
template <int NCMAX>
__host__ __device__ __forceinline__ void word_classify_met34(uint32_t own, const uint32_t (&nb)[4],
                                                             uint32_t valid, const ClassTab<NCMAX>& t,
                                                             WordDecision<NCMAX>& w)
{
    uint32_t in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) in[k] = ~(own ^ nb[k]);  // aligned neighbours
    count_bits<4>(in, w.c0, w.c1, w.c2);
#pragma unroll
    for (int k = 0; k < NCMAX; ++k) {
        w.mask[k] = k == 0 ? w.c0 & w.c1 : k == 1 ? w.c2 : 0u;
        w.thr[k] = t.thr[k];
    }
    const uint32_t cls = w.mask[0] | w.mask[1];
    w.pending = cls & valid;
    w.below = valid & ~cls;
}

// True when a decoded Metropolis table is the z = 4, zero-field one that
// word_classify_met34 handles: two classes, aligned counts 3 and 4, no spin
// selection.
__host__ __device__ __forceinline__ bool table_is_met34(const uint32_t* __restrict__ tab)
{
    return tab[0] == 2u && tab[1] == 3u && tab[2] == 4u;
}

// Continue the comparison with Philox blocks q_begin .. q_end-1 of the word's
// group (block 8 group + q gives planes 4q .. 4q+3, most significant first);
// stops early once no lane is pending, which cannot change a decision.
template <int NCMAX>
__host__ __device__ __forceinline__ void word_draw(WordDecision<NCMAX>& w, uint32_t q_begin,
                                                   uint32_t q_end, uint32_t k0, uint32_t k1,
                                                   uint32_t sid, uint64_t sweep, uint32_t colour,
                                                   uint32_t group)
{
    const uint32_t zc = ctr_z(sweep, colour, TAG_ISING);
    for (uint32_t q = q_begin; q < q_end && w.pending != 0u; ++q) {
        const u32x4 r = philox4x32_10(u32x4{8u * group + q, (uint32_t)sweep, zc, sid}, k0, k1);
        const uint32_t planes[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t shift = 31u - (4u * q + (uint32_t)t);
            // class k's threshold bit as an all-ones / all-zeros mask, uniform
            // across the wave: tb takes one AND-OR per class
            uint32_t tb = 0;
#pragma unroll
            for (int k = 0; k < NCMAX; ++k)
                tb |= w.mask[k] & (uint32_t)((int32_t)(w.thr[k] << (31u - shift)) >> 31);
            const uint32_t diff = w.pending & (planes[t] ^ tb);
            w.below |= diff & tb;
            w.pending &= ~diff;
        }
    }
}

// Apply a finished decision: returns the new word and accumulates deltas.
template <int Z, int NCMAX>
__host__ __device__ __forceinline__ uint32_t word_apply(int code, uint32_t own, uint32_t valid,
                                                        const WordDecision<NCMAX>& w, Delta& d)
{
    if (code == CODE_METROPOLIS) {
        const uint32_t F = w.below & valid;
        const int nf = popc(F);
        const int nal = popc(F & w.c0) + 2 * popc(F & w.c1) + 4 * popc(F & w.c2);
        d.de += 4 * nal - 2 * Z * nf;          // sum of 2a, a = 2 n_aligned - z
        d.dm -= 2 * (2 * popc(F & own) - nf);  // sum of -2 s_old
        d.acc += nf;
        return own ^ F;
    }
    const uint32_t X = (w.below ^ own) & valid;
    const uint32_t Xu = X & own, Xd = X & ~own;
    const int fu = 2 * (popc(Xu & w.c0) + 2 * popc(Xu & w.c1) + 4 * popc(Xu & w.c2)) - Z * popc(Xu);
    const int fd = 2 * (popc(Xd & w.c0) + 2 * popc(Xd & w.c1) + 4 * popc(Xd & w.c2)) - Z * popc(Xd);
    d.de += 2 * (fu - fd);                 // sum of 2 s_old f
    d.dm += 2 * (popc(Xd) - popc(Xu));     // sum of 2 s_new
    d.acc += popc(X);
    return own ^ X;
}

// Metropolis flips counted raw, per lane, and folded into a Delta once per
// measurement (word_apply's deltas are linear in these counts): n_b =
// popc(F & c_b) for the count bits, nf = popc(F), no = popc(F & own), so
// de = 4 (n0 + 2 n1 + 4 n2) - 2 Z nf and dm = -2 (2 no - nf).
struct FlipCounts {
    int n0, n1, n2, nf, no;
};

template <int NCMAX>
__host__ __device__ __forceinline__ uint32_t word_apply_metropolis(uint32_t own, uint32_t valid,
                                                                   const WordDecision<NCMAX>& w,
                                                                   FlipCounts& f)
{
    const uint32_t F = w.below & valid;
    f.nf += popc(F);
    f.n0 += popc(F & w.c0);
    f.n1 += popc(F & w.c1);
    f.n2 += popc(F & w.c2);
    f.no += popc(F & own);
    return own ^ F;
}

template <int Z>
__host__ __device__ __forceinline__ void fold_counts(FlipCounts& f, Delta& d)
{
    d.de += 4 * (f.n0 + 2 * f.n1 + 4 * f.n2) - 2 * Z * f.nf;
    d.dm -= 2 * (2 * f.no - f.nf);
    d.acc += f.nf;
    f = FlipCounts{0, 0, 0, 0, 0};
}

// One word of 32 same-colour sites; returns the new spins and accumulates
// deltas.  "u < p" for the site at bit b is "x < T" with x the 32-bit uniform
// whose bit 31-j is bit b of plane j; the comparison runs plane by plane.
template <int Z, int NCMAX>
__host__ __device__ __forceinline__ uint32_t update_word(int code, uint32_t own, const uint32_t (&nb)[Z],
                                                uint32_t valid, const uint32_t* __restrict__ tabs,
                                                uint32_t k0, uint32_t k1, uint32_t sid,
                                                uint64_t sweep, uint32_t colour, uint32_t group,
                                                Delta& d)
{
    WordDecision<NCMAX> w;
    word_classify<Z, NCMAX>(code, own, nb, valid, tabs, w);
    word_draw<NCMAX>(w, 0, 8, k0, k1, sid, sweep, colour, group);
    return word_apply<Z, NCMAX>(code, own, valid, w, d);
}

}  // namespace cm
