// Classical O(n) vector model (XY n = 2, Heisenberg n = 3): FP32 spins,
// checkerboard Metropolis and over-relaxation sweeps, FP64 renormalisation
// and measurements.
//
// Precision: spins and the proposal / reflection arithmetic are FP32; the
// local field is accumulated in FP64 and rounded once (see on_half_sweep).
//
// Reference (critmc/kernels/on.py):
//   Metropolis   :112-137  S' = (S + step g)/|S + step g|, g ~ N(0,1)^n;
//                          dE = -(S' - S).h; accept iff dE <= 0 or u < exp(-beta dE)
//   over-relax   :140-183  S' = 2 (S.h)/|h|^2 h - S, unchanged when |h|^2 = 0
//   measurement  :35-87, 404-413  normalise, then E = -sum S_i.S_j, M, and per
//                axis c_a = sum (S_i^0 S_j^0 + S_i^1 S_j^1), s_a = sum (S_i^0 S_j^1 - S_i^1 S_j^0)
//
// Layout: spins [R][2 colours][N/2][n] FP32: site i (x fastest) of colour
// c = (x0 + x1 + x2) & 1 sits in plane c at i >> 1 (cm_on_convert translates
// from and to site order).  A half sweep then reads and writes its own plane
// contiguously, and its neighbours are all in the other plane at nb >> 1:
// full cache lines instead of every other 8 bytes.  A half sweep is one launch
// over all replicas (grid = colour-site chunks x replicas).  Two consecutive
// over-relaxation sweeps of XY on a square lattice run as one launch over row
// tiles in LDS (on_or_fused), bit-identical to four half-sweep launches.
// Lattices whose spins fit in ON_BATCHED_LDS run a call of local updates as
// one launch instead (on_batched: one workgroup per replica, spins in LDS,
// the same per-site functions, so the same bits).
// When the schedule ends with a local update, that last half sweep writes its
// sites renormalised and colour 0 is renormalised while every bond is summed
// once: by the next measurement's first colour-0 half sweep, which then
// updates those sites (on_c0_sweep_measure), or by on_measure_c0 for a call's last
// measurement (when it ends with two fused over-relaxation sweeps, the fused
// launch does both); otherwise a normalise launch precedes the partial-sum
// launch.  Partial sums are finalised in order, so every FP64 result is
// reproducible run to run.
//
// Heat bath (:186-261), FP64 from the exact FP64 field:
//   XY         phi = atan2(h) + von Mises(kappa = beta |h|) by Best & Fisher's
//              rejection; attempt a of colour site k uses Philox block k with
//              the attempt in bits 1..7 of the colour field (tag TAG_ONHB):
//              word 0 -> z = cos(pi U), word 1 -> u2 in (0, 1], word 2's top bit
//              -> sign.  After 128 rejections the last proposal is kept
//              (probability < 1e-50).  kappa < 1e-8: phi uniform from word 0.
//   Heisenberg u = 1 + ln(r + (1 - r) e^{-2 kappa}) / kappa along h, azimuth
//              2 pi U in the reference's (h, e1, e2) frame; block k words 0, 1;
//              |h| = 0: z = 2U - 1 and azimuth from words 2, 3.
//
// Cluster updates (:264-345), the embedded Ising model along a random unit
// vector r, uniforms at tag TAG_CLUSTER (one sweep-counter value per update):
//   r: block 0 of colour field 3 -- XY angle 2 pi U (word 0); n = 3:
//      z = 2U - 1 (word 0), azimuth 2 pi U (word 1);
//   p_i = S_i . r in FP64; bond b = a N + i (i and its +e_a neighbour j) is
//   active iff x = p_i p_j > 0 and u < 1 - det_exp(-2 beta x), u = word b&3 of
//   block b>>2 (colour field 0); det_exp is the fixed-operation e^x, so the
//   host replay takes the same decisions;
//   Swendsen-Wang: the cluster whose smallest site is rt reflects,
//   S -> S - 2 p r, iff word rt&3 of block rt>>2 (colour field 1) has its top
//   bit clear; Wolff: seed floor(x N / 2^32) (word 0 of block 0, colour field
//   2), a neighbour j of member k joins iff its bond is active with p_k the
//   member's projection before its reflection (the reference's order of
//   effects; every bond has its own uniform, so the cluster does not depend on
//   the order of exploration).  Reflections are FP64, stored FP32.
//
// Uniforms (Metropolis, tag TAG_ON): colour site k of a replica uses Philox
// block k (n = 2: words 0,1 -> Box-Muller pair, word 2 -> acceptance) or blocks
// 2k, 2k+1 (n = 3: words 0..3 -> two Box-Muller pairs, third Gaussian from the
// first of the second pair; block 2k+1 word 0 -> acceptance).  u1 = (x+1)/2^32
// in (0, 1], u2 = x/2^32 in [0, 1).
#include <algorithm>
#include <cmath>
#include <vector>

#include "cb_geom.hpp"
#include "cluster_core.hpp"
#include "common.hpp"
#include "det_exp.hpp"
#include "ising_core.hpp"
#include "philox.hpp"

namespace cm {

constexpr int ON_CHUNK = 4096;  // sites per measurement partial

template <int NC>
struct Vec {
    float v[NC];
};

template <int NC>
__device__ __forceinline__ Vec<NC> load_vec(const float* __restrict__ p, int64_t i)
{
    Vec<NC> r;
    if constexpr (NC == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p + 2 * i);
        r.v[0] = t.x;
        r.v[1] = t.y;
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) r.v[c] = p[NC * i + c];
    }
    return r;
}

template <int NC>
__device__ __forceinline__ void store_vec(float* __restrict__ p, int64_t i, const Vec<NC>& r)
{
    if constexpr (NC == 2) {
        *reinterpret_cast<float2*>(p + 2 * i) = make_float2(r.v[0], r.v[1]);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) p[NC * i + c] = r.v[c];
    }
}

// Renormalise in FP64 (critmc/kernels/on.py:78-87), rounded to FP32.  The
// squares of the FP32 components are exact in FP64.  Near |S| = 1 (every spin
// after a sweep: |S|^2 - 1 is of order 1e-6), 1/|S| is the series
// 1 - d/2 + 3d^2/8 - 5d^3/16 + 35d^4/128 in d = |S|^2 - 1, whose remainder is
// below 2^-60 for |d| <= 2^-12; farther from 1, the FP64 square root and
// division.  Every renormalisation on the device goes through this function,
// so fused and separate launches store the same bits.
template <int NC>
__device__ __forceinline__ Vec<NC> unit_vec(Vec<NC> s)
{
    double n2 = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) n2 += (double)s.v[c] * (double)s.v[c];
    const double d = n2 - 1.0;
    if (fabs(d) <= 0x1p-12) {
        const double inv = fma(d, fma(d, fma(d, fma(d, 35.0 / 128.0, -5.0 / 16.0), 3.0 / 8.0), -0.5), 1.0);
#pragma unroll
        for (int c = 0; c < NC; ++c) s.v[c] = (float)((double)s.v[c] * inv);
    } else {
        const double nrm = sqrt(n2);
#pragma unroll
        for (int c = 0; c < NC; ++c) s.v[c] = (float)((double)s.v[c] / nrm);
    }
    return s;
}

__device__ __forceinline__ float2 unit_xy(float2 v)
{
    Vec<2> t{{v.x, v.y}};
    t = unit_vec<2>(t);
    return make_float2(t.v[0], t.v[1]);
}

// Over-relaxation S' = f h - S with f = 2 (S.h) / |h|^2, S unchanged when
// |h|^2 = 0 (critmc/kernels/on.py:140-183).  FP32; 1/|h|^2 from v_rcp_f32
// (1 ulp) for a normal |h|^2, the IEEE division below that.  Both the
// per-half-sweep kernel and the fused LDS kernel use it, so they agree bit
// for bit.
template <int NC>
__device__ __forceinline__ bool reflect(Vec<NC>& s, const float (&h)[NC])
{
    float hh = h[0] * h[0], sh = s.v[0] * h[0];
#pragma unroll
    for (int c = 1; c < NC; ++c) {
        hh += h[c] * h[c];
        sh += s.v[c] * h[c];
    }
    if (hh == 0.0f) return false;
    float f;
    if (hh >= 0x1p-126f) f = 2.0f * sh * __builtin_amdgcn_rcpf(hh);
    else f = 2.0f * sh / hh;
#pragma unroll
    for (int c = 0; c < NC; ++c) s.v[c] = f * h[c] - s.v[c];
    return true;
}

// Box-Muller pair from two Philox words: radius sqrt(-2 ln u1), u1 = (xa+1)/2^32
// in (0, 1], from v_log_f32 (log2) and v_sqrt_f32, each within about 1 ulp;
// angle 2 pi xb / 2^32 split into the nearest quadrant q (top bits of xb) and
// a remainder theta in [-pi/4, pi/4) taken exactly from the low 30 bits, then
// Taylor polynomials for sin and cos (remainders below 2e-9 on that range)
// and a rotation by q quadrants.  Largest error against the exact angle
// 2 pi xb / 2^32: 1e-7 over 2^24 angles, all quadrant edges included.
__device__ __forceinline__ void box_muller(uint32_t xa, uint32_t xb, float& g0, float& g1)
{
    const float u1 = ((float)xa + 1.0f) * 0x1p-32f;
    const float r = __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u1));
    const uint32_t sh = xb + 0x20000000u;
    const uint32_t q = sh >> 30;
    const int32_t rem = (int32_t)(sh & 0x3FFFFFFFu) - 0x20000000;
    const float th = (float)rem * (6.28318530717958647692f * 0x1p-32f);
    const float t2 = th * th;
    const float ps = fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f),
                                   -1.0f / 6.0f), 0.0f);
    const float sn = fmaf(th, ps, th);
    const float cs = fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, -1.0f / 3628800.0f, 1.0f / 40320.0f),
                                                          -1.0f / 720.0f), 1.0f / 24.0f), -0.5f), 1.0f);
    float c = (q & 1) ? -sn : cs, s = (q & 1) ? cs : sn;
    if (q & 2) {
        c = -c;
        s = -s;
    }
    g0 = r * c;
    g1 = r * s;
}

__device__ __forceinline__ double unit_open0(uint32_t x) { return ((double)x + 1.0) * 0x1p-32; }
__device__ __forceinline__ double unit_open1(uint32_t x) { return (double)x * 0x1p-32; }

// Best & Fisher (1979) von Mises deviate theta (critmc/kernels/on.py:186-204),
// given as its cosine and sine: the rejection loop ends with cos(theta) = f,
// and the deviate is +-acos(f), so sin(theta) = +-sqrt(1 - f^2) with the
// loop's sign.  The caller rotates (ct, st) by the field's direction instead
// of forming atan2(h) + theta and taking its cosine and sine; the decisions,
// and the uniforms each attempt draws, are those of the FP64 rule.  Returns
// false for kappa < 1e-8, where the deviate is uniform: `angle` then holds it.
__device__ bool von_mises_cs(double kappa, uint32_t blk, uint64_t sweep, uint32_t colour,
                             uint32_t sid, uint32_t k0, uint32_t k1, double& ct, double& st,
                             double& angle)
{
    if (kappa < 1e-8) {
        const u32x4 b = philox4x32_10(u32x4{blk, (uint32_t)sweep, ctr_z(sweep, colour, TAG_ONHB), sid}, k0, k1);
        angle = M_PI * (2.0 * unit_open1(b.x) - 1.0);
        return false;
    }
    const double root = sqrt(1.0 + 4.0 * kappa * kappa);
    const double tau = 1.0 + root;
    const double rho = 2.0 * kappa * tau / ((root + 1.0) * (tau + sqrt(2.0 * tau)));
    const double r = (1.0 + rho * rho) / (2.0 * rho);
    double f = 0.0;
    uint32_t sign = 0;
    for (uint32_t a = 0; a < 128; ++a) {
        const u32x4 b = philox4x32_10(
            u32x4{blk, (uint32_t)sweep, ctr_z(sweep, colour | (a << 1), TAG_ONHB), sid}, k0, k1);
        const double z = cos(M_PI * unit_open1(b.x));
        f = (1.0 + r * z) / (r + z);
        const double c = kappa * (r - f);
        const double u2 = unit_open0(b.y);
        sign = b.z >> 31;
        if (c * (2.0 - c) - u2 > 0.0 || log(c / u2) + 1.0 - c >= 0.0) break;
    }
    ct = fmin(1.0, fmax(-1.0, f));
    const double sn = sqrt(fmax(0.0, 1.0 - ct * ct));  // sin(acos(ct)) >= 0
    st = sign == 0 ? sn : -sn;
    return true;
}

struct OnArgs {
    float* spins;
    const float* betas;     // [R]
    const float* steps;     // [R]
    const uint32_t* keys;   // [R][2]
    const uint32_t* sids;   // [R]
    const uint64_t* sweeps; // [R]
    int32_t* acc;           // [R][n_meas] (accumulated)
    const int64_t* n_meas_rep;  // [R] measurements per replica, or null
    CbGeo g;
    int R;
    int64_t n_meas, meas;
    uint64_t sweep_add;
    int colour;
    int norm_out;  // write every site of this half sweep renormalised (the measurement's last)
};

// The update of one site of colour `colour` (index k within its colour), the
// per-site work every local O(n) update shares: the per-half-sweep kernels
// and the per-replica LDS kernel both call it, so they store the same bits.
// `So` is the other colour's plane (HBM or LDS) and nbo the neighbours'
// offsets in it; s is the site's spin, updated in place.  Returns whether s
// changed; `accepted` is set to 1 when a Metropolis proposal is accepted.
template <int NC, int Z, int CODE>
__device__ __forceinline__ bool on_site(Vec<NC>& s, const float* __restrict__ So,
                                        const uint32_t (&nbo)[Z], uint32_t k, int colour,
                                        uint64_t sweep, uint32_t k0, uint32_t k1, uint32_t sid,
                                        float beta, float step, int& accepted)
{
    // The local field is summed in FP64: its Z FP32 terms are exact there, so
    // a nearly cancelling field keeps full relative precision (in FP32 the
    // sum's absolute error eps * Z would dominate when |h| -> 0).
    double hd[NC];
#pragma unroll
    for (int j = 0; j < Z; ++j) {
        const Vec<NC> t = load_vec<NC>(So, nbo[j]);
#pragma unroll
        for (int c = 0; c < NC; ++c) hd[c] = j == 0 ? (double)t.v[c] : hd[c] + (double)t.v[c];
    }
    float h[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) h[c] = (float)hd[c];
    bool changed = false;
    if (CODE == CODE_HEATBATH) {
        double hn2 = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) hn2 += hd[c] * hd[c];
        const double hn = sqrt(hn2);
        const double kappa = (double)beta * hn;
        if constexpr (NC == 2) {
            double ct, st, theta;
            if (von_mises_cs(kappa, k, sweep, (uint32_t)colour, sid, k0, k1, ct, st, theta)) {
                // new angle phi0 + theta with phi0 the field's direction h / |h|
                // (kappa >= 1e-8 implies |h| > 0)
                const double inv = 1.0 / hn;
                const double cx = hd[0] * inv, sy = hd[1] * inv;
                s.v[0] = (float)(cx * ct - sy * st);
                s.v[1] = (float)(sy * ct + cx * st);
            } else {
                const double phi = (hn > 0.0 ? atan2(hd[1], hd[0]) : 0.0) + theta;
                s.v[0] = (float)cos(phi);
                s.v[1] = (float)sin(phi);
            }
        } else {
            const u32x4 b = philox4x32_10(
                u32x4{k, (uint32_t)sweep, ctr_z(sweep, (uint32_t)colour, TAG_ONHB), sid}, k0, k1);
            if (hn == 0.0) {
                const double z = 2.0 * unit_open1(b.z) - 1.0;
                const double az = 2.0 * M_PI * unit_open1(b.w);
                const double w = sqrt(fmax(0.0, 1.0 - z * z));
                s.v[0] = (float)(w * cos(az));
                s.v[1] = (float)(w * sin(az));
                s.v[2] = (float)z;
            } else {
                const double r = unit_open0(b.x);
                double u = kappa < 1e-8 ? 2.0 * r - 1.0
                                        : 1.0 + log(r + (1.0 - r) * exp(-2.0 * kappa)) / kappa;
                u = fmin(1.0, fmax(-1.0, u));
                const double hx = hd[0] / hn, hy = hd[1] / hn, hz = hd[2] / hn;
                const double ax = fabs(hx) < 0.9 ? 1.0 : 0.0, ay = 1.0 - ax;
                const double proj = ax * hx + ay * hy;
                double e1x = ax - proj * hx, e1y = ay - proj * hy, e1z = -proj * hz;
                const double nrm = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
                e1x /= nrm;
                e1y /= nrm;
                e1z /= nrm;
                const double e2x = hy * e1z - hz * e1y, e2y = hz * e1x - hx * e1z,
                             e2z = hx * e1y - hy * e1x;
                const double phi = 2.0 * M_PI * unit_open1(b.y);
                const double w = sqrt(fmax(0.0, 1.0 - u * u));
                const double cp = cos(phi), sp = sin(phi);
                s.v[0] = (float)(u * hx + w * (cp * e1x + sp * e2x));
                s.v[1] = (float)(u * hy + w * (cp * e1y + sp * e2y));
                s.v[2] = (float)(u * hz + w * (cp * e1z + sp * e2z));
            }
        }
        changed = true;
    } else if (CODE == CODE_OVERRELAX) {
        changed = reflect<NC>(s, h);
    } else {
        const uint32_t zc = ctr_z(sweep, (uint32_t)colour, TAG_ON);
        float g[4];
        uint32_t xacc;
        if constexpr (NC == 2) {
            const u32x4 r = philox4x32_10(u32x4{k, (uint32_t)sweep, zc, sid}, k0, k1);
            box_muller(r.x, r.y, g[0], g[1]);
            xacc = r.z;
        } else {
            const u32x4 r = philox4x32_10(u32x4{2u * k, (uint32_t)sweep, zc, sid}, k0, k1);
            const u32x4 r2 = philox4x32_10(u32x4{2u * k + 1u, (uint32_t)sweep, zc, sid}, k0, k1);
            box_muller(r.x, r.y, g[0], g[1]);
            box_muller(r.z, r.w, g[2], g[3]);
            xacc = r2.x;
        }
        float tr[NC];
        float n2 = 0.0f;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            tr[c] = s.v[c] + step * g[c];
            n2 += tr[c] * tr[c];
        }
        // v_rsq_f32 (1 ulp) scales the proposal: its direction is exact, its
        // length is 1 within 2e-7 and renormalised at every measurement
        const float inv = __builtin_amdgcn_rsqf(n2);
        float de = 0.0f;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            tr[c] *= inv;
            de -= (tr[c] - s.v[c]) * h[c];
        }
        // v_exp_f32: a decision can differ from the FP64 rule only where u lies
        // within a few ulp of the acceptance probability
        if (de <= 0.0f || (float)xacc * 0x1p-32f < __expf(-beta * de)) {
#pragma unroll
            for (int c = 0; c < NC; ++c) s.v[c] = tr[c];
            changed = true;
            accepted = 1;
        }
    }
    return changed;
}

template <int NC, int D>
__device__ __forceinline__ void on_c0_bonds(const Vec<NC>& s, const float* __restrict__ S1,
                                            const uint32_t (&nbo)[2 * D], double (&v)[1 + NC + 2 * D]);

// The ordered sum of one quantity of one replica's colour-0 block partials,
// taken by one wave (pair = replica * K + quantity, wave-uniform) and written
// to its output series: the lanes load 512 partials at a time in parallel and
// the wave adds them in block order from lane broadcasts, so the sum is the
// plain left-to-right one without a chain of dependent loads.
template <int NC, int D>
__device__ __forceinline__ void on_final_pair(const double* __restrict__ part, int nb, int R,
                                              int64_t n_meas, int64_t meas, double* __restrict__ e_out,
                                              double* __restrict__ m_out, double* __restrict__ sc_out,
                                              double* __restrict__ ss_out,
                                              const int64_t* __restrict__ n_meas_rep, int pair)
{
    constexpr int K = 1 + NC + 2 * D;
    if (pair >= R * K) return;
    const int rep = pair / K, q = pair - (pair / K) * K;
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;
    const int lane = (int)(threadIdx.x & 63);
    const double* pp = part + (int64_t)rep * nb * K + q;
    double s = 0.0;
    for (int c0 = 0; c0 < nb; c0 += 512) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = c0 + 64 * j + lane;
            v[j] = c < nb ? pp[(int64_t)c * K] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int lim = nb - c0 - 64 * j;  // partials of this group of 64 (wave-uniform)
#pragma unroll
            for (int l = 0; l < 64; ++l) {
                if (l >= lim) break;
                const uint64_t bits = (uint64_t)__double_as_longlong(v[j]);
                const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bits, l);
                const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), l);
                s += __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
            }
        }
    }
    if (lane != 0) return;
    const int64_t o = (int64_t)rep * n_meas + meas;
    if (q == 0) e_out[o] = s;
    else if (q < 1 + NC) m_out[o * NC + (q - 1)] = s;
    else if (q < 1 + NC + D) sc_out[o * D + (q - 1 - NC)] = s;
    else ss_out[o * D + (q - 1 - NC - D)] = s;
}

// A 256-lane block's partial of K quantities: each wave's sum, then the four
// waves added in order (thread q < K writes quantity q to dst[q]).
template <int K>
__device__ __forceinline__ void on_block_partial(const double (&v)[K], double* __restrict__ red,
                                                 double* __restrict__ dst)
{
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const double t = wave_sum_f64(v[q]);
        if ((threadIdx.x & 63) == 0) red[q * 4 + w] = t;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double t = 0.0;
        for (int u = 0; u < 4; ++u) t += red[threadIdx.x * 4 + u];
        dst[threadIdx.x] = t;
    }
}

// The ordered sums of a measurement whose chunk partials an earlier launch
// formed (on_c0_sweep_measure), taken by the next half-sweep launch in an
// extra row of workgroups: one wave per (replica, quantity).
struct OnFinal {
    const double* part;  // null: none
    int64_t meas;
    int nchunk;          // partials per replica
    double* e_out;
    double* m_out;
    double* sc_out;
    double* ss_out;
};

template <int NC, int Z, int CODE, bool FIN>
__global__ __launch_bounds__(256) void on_half_sweep(OnArgs a, OnFinal fin)
{
    constexpr int D = Z / 2, K = 1 + NC + Z;
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x);
    const uint32_t bx = lb - (uint32_t)rep * gridDim.x;
    if constexpr (FIN) {
        if (rep == a.R) {  // the extra row of workgroups
            const int waves = (int)(gridDim.x * (blockDim.x >> 6));
            for (int pr = (int)(bx * (blockDim.x >> 6) + (threadIdx.x >> 6)); pr < a.R * K; pr += waves)
                on_final_pair<NC, D>(fin.part, fin.nchunk, a.R, a.n_meas, fin.meas, fin.e_out, fin.m_out,
                                     fin.sc_out, fin.ss_out, a.n_meas_rep, pr);
            return;
        }
    }
    if (a.n_meas_rep && a.meas >= a.n_meas_rep[rep]) return;  // whole block: replica done
    const int64_t k = (int64_t)bx * blockDim.x + threadIdx.x;
    int accepted = 0;
    if (k < a.g.half) {
        float* S = a.spins + (int64_t)rep * a.g.N * NC;
        uint32_t nbo[Z];
        cb_other<Z>(a.g, (uint32_t)k, a.colour, nbo);
        const int64_t i = (int64_t)a.colour * a.g.half + k;            // own plane
        const float* So = S + (int64_t)(1 - a.colour) * a.g.half * NC;  // neighbours' plane
        Vec<NC> s = load_vec<NC>(S, i);
        const bool changed = on_site<NC, Z, CODE>(s, So, nbo, (uint32_t)k, a.colour,
                                                  a.sweeps[rep] + a.sweep_add, a.keys[2 * rep],
                                                  a.keys[2 * rep + 1], a.sids[rep], a.betas[rep],
                                                  a.steps[rep], accepted);
        if (a.norm_out) {
            // the measurement's renormalisation of this site, fused into its last update
            s = unit_vec<NC>(s);
            store_vec<NC>(S, i, s);
        } else if (changed) {
            store_vec<NC>(S, i, s);
        }
    }
    if (CODE == CODE_METROPOLIS) {  // one atomic per workgroup
        __shared__ int part[4];
        const int t = wave_sum_bits<1>((uint32_t)accepted);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int tot = part[0] + part[1] + part[2] + part[3];
            if (tot) atomicAdd(&a.acc[(int64_t)rep * a.n_meas + a.meas], tot);
        }
    }
}

// Over-relaxation sweeps of XY spins on a square lattice, K half sweeps
// (colour 0, 1, 0, 1, ...) fused in LDS.  A workgroup owns rows
// [y0, y0 + T) of one replica: it loads colour 0 on rows y0-(K-1) .. y0+T+K-2
// and colour 1 on rows y0-K .. y0+T+K-1, runs the half sweeps on shrinking row
// ranges (half sweep h covers the rows still needed by the later ones), and
// writes back its own rows.  Every site update does the same operations in
// the same order as on_half_sweep, so the result is bit-identical to K
// separate launches, with about a third of their HBM traffic.
//
// With `part` set (the measurement's last launch, a.norm_out), the workgroup
// also measures: colour 0 is final on rows y0-1 .. y0+T after half sweep K-2,
// so those rows are renormalised in LDS (the values on_measure_c0 would
// store), and each colour-1 site of the own rows takes its 2d bonds and M of
// itself and of its -e_0 neighbour (a bijection onto colour 0).  The tile's
// FP64 partial sums go to part[rep][tile][7]; on_measure_final adds the
// tiles in order.
template <int K>
__global__ __launch_bounds__(1024) void on_or_fused(OnArgs a, const float* __restrict__ src, int T,
                                                    double* __restrict__ part)
{
    extern __shared__ float2 orl[];
    const int M = a.g.M, L1 = a.g.L1;
    const int tiles = (L1 + T - 1) / T;
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int rep = (int)(lb / (uint32_t)tiles), tile = (int)(lb - (uint32_t)rep * tiles);
    const int y0 = tile * T;
    const int Te = min(T, L1 - y0);
    const int R0 = Te + 2 * (K - 1), R1 = Te + 2 * K;
    float2* P0 = orl;
    float2* P1 = orl + (size_t)R0 * M;
    // out of place: the tile reads its rows and halos from `src` and writes its
    // own rows to a.spins, so no tile can read a row another has already updated
    const float2* I0 = (const float2*)(src + (int64_t)rep * a.g.N * 2);
    const float2* I1 = I0 + a.g.half;
    float2* G0 = (float2*)(a.spins + (int64_t)rep * a.g.N * 2);
    float2* G1 = G0 + a.g.half;
    // Each thread keeps one column mm and visits rows tr0, tr0 + bq, ...: the
    // workgroup covers bq whole rows per pass (threads beyond bq rows idle),
    // so LDS offsets advance by a constant and no per-site division is left.
    const int bq = (int)fdiv(blockDim.x, a.g.divM);
    const int tr0 = (int)fdiv(threadIdx.x, a.g.divM);
    const int mm = (int)threadIdx.x - tr0 * M;
    const bool act = tr0 < bq;
    const int stride = bq * M;
    const int dp1 = mm == M - 1 ? -(M - 1) : 1;  // x0 + 1 in the other plane, x0 odd
    const int dn0 = mm == 0 ? M - 1 : -1;        // x0 - 1 in the other plane, x0 even
    if (a.n_meas_rep && a.n_meas_rep[rep] <= 0) return;  // no measurement this call: state stays put
    if (a.n_meas_rep && a.meas >= a.n_meas_rep[rep]) {  // whole block: replica done, rows carried over
        if (act)
            for (int r = tr0; r < Te; r += bq) {
                G0[(int64_t)(y0 + r) * M + mm] = I0[(int64_t)(y0 + r) * M + mm];
                G1[(int64_t)(y0 + r) * M + mm] = I1[(int64_t)(y0 + r) * M + mm];
            }
        return;
    }
    // lattice row of LDS row r (the halo rows wrap; at most once when L1 >= K)
    auto wrap = [&](int y) {
        if (L1 >= K) return y < 0 ? y + L1 : (y >= L1 ? y - L1 : y);
        const int q = y % L1;
        return q < 0 ? q + L1 : q;
    };
    if (act) {
        for (int r = tr0; r < R0; r += bq) P0[r * M + mm] = I0[(int64_t)wrap(y0 - (K - 1) + r) * M + mm];
        for (int r = tr0; r < R1; r += bq) P1[r * M + mm] = I1[(int64_t)wrap(y0 - K + r) * M + mm];
    }
    __syncthreads();
#pragma unroll
    for (int hs = 0; hs < K; ++hs) {
        const int c = hs & 1;
        const int ext = K - 1 - hs;  // rows beyond the tile still needed later
        const int nrows = Te + 2 * ext;
        float2* Pc = c ? P1 : P0;
        const float2* Po = c ? P0 : P1;
        const int offc = c ? K : K - 1, offo = c ? K - 1 : K;  // local row of y0 in Pc / Po
        const bool norm = a.norm_out && hs == K - 1;
        if (act) {
            int io = (tr0 - ext + offo) * M + mm;  // this site's row and column in Po
            const int dc = (offc - offo) * M;      // ... and in Pc
            const int pb = y0 - ext + c;           // L1 is even, so parity needs no wrap
            for (int rr = tr0; rr < nrows; rr += bq, io += stride) {
                const bool odd = ((rr + pb) & 1) != 0;  // own site at x0 = 2 mm + 1
                const float2 nb[4] = {Po[io + (odd ? dp1 : 0)], Po[io + (odd ? 0 : dn0)], Po[io + M],
                                      Po[io - M]};
                double hd0 = (double)nb[0].x, hd1 = (double)nb[0].y;
#pragma unroll
                for (int j = 1; j < 4; ++j) {
                    hd0 += (double)nb[j].x;
                    hd1 += (double)nb[j].y;
                }
                const float h[2] = {(float)hd0, (float)hd1};
                const float2 v0 = Pc[io + dc];
                Vec<2> v{{v0.x, v0.y}};
                (void)reflect<2>(v, h);
                float2 out = make_float2(v.v[0], v.v[1]);
                if (norm) out = unit_xy(out);
                Pc[io + dc] = out;
            }
        }
        __syncthreads();
    }
    if (part) {
        float2* Q = P0 + (K - 2) * M;  // colour 0 from row y0 - 1
        if (act)
            for (int r = tr0; r < Te + 2; r += bq) Q[r * M + mm] = unit_xy(Q[r * M + mm]);
        __syncthreads();
        double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // E, M[2], c_a[2], s_a[2]
        if (act) {
            int io = (K - 1 + tr0) * M + mm;  // P0 row and column of the site; P1 is one row on
            for (int rr = tr0; rr < Te; rr += bq, io += stride) {
                const bool odd = ((y0 + rr + 1) & 1) != 0;  // colour-1 site at x0 = 2 mm + 1
                const float2 s = P1[io + M];
                // forward (j + e_a) and backward (j - e_a) neighbours per axis
                const float2 fw[2] = {P0[io + (odd ? dp1 : 0)], P0[io + M]};
                const float2 bw[2] = {P0[io + (odd ? 0 : dn0)], P0[io - M]};
                const double sx = s.x, sy = s.y;
                acc[1] += sx + (double)bw[0].x;
                acc[2] += sy + (double)bw[0].y;
                double cs = 0.0;
#pragma unroll
                for (int ax = 0; ax < 2; ++ax) {
                    // bonds (j, j + e_a) and (j - e_a, j), oriented as the reference's
                    const double fx = fw[ax].x, fy = fw[ax].y, bx = bw[ax].x, by = bw[ax].y;
                    const double cax = sx * (fx + bx) + sy * (fy + by);
                    acc[3 + ax] += cax;
                    acc[5 + ax] += sx * (fy - by) - sy * (fx - bx);
                    cs += cax;
                }
                acc[0] -= cs;
            }
        }
        __shared__ double red[7 * 16];
        const int w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const double t = wave_sum_f64(acc[q]);
            if ((threadIdx.x & 63) == 0) red[q * 16 + w] = t;
        }
        __syncthreads();
        if (threadIdx.x < 7) {
            double t = 0.0;
            for (int u = 0; u < nw; ++u) t += red[threadIdx.x * 16 + u];
            part[((int64_t)rep * tiles + tile) * 7 + threadIdx.x] = t;
        }
    }
    if (act)
        for (int r = tr0; r < Te; r += bq) {
            G0[(int64_t)(y0 + r) * M + mm] = P0[(K - 1 + r) * M + mm];
            G1[(int64_t)(y0 + r) * M + mm] = P1[(K + r) * M + mm];
        }
}

constexpr int ON_OR_FUSE = 4;                // half sweeps per fused launch (two sweeps)
#ifndef CM_ON_OR_LDS_KB
#define CM_ON_OR_LDS_KB 76
#endif
constexpr int64_t ON_OR_LDS = CM_ON_OR_LDS_KB * 1024;  // LDS per workgroup (76 KB: two per CU)

// Rows per tile for the fused over-relaxation (0: the lattice does not qualify).
static int on_or_tile_rows(const CbGeo& g, int n)
{
    if (n != 2 || g.d != 2 || g.M < 2) return 0;
    const int64_t row = (int64_t)g.M * 8;
    const int64_t T = (ON_OR_LDS / row - (4 * ON_OR_FUSE - 2)) / 2;
    if (T < 4) return 0;
    return (int)std::min<int64_t>(T, g.L1);
}

// The spins of every replica that made a measurement in this call, from the
// buffer the last row-tile launch wrote back to the caller's (replicas with no
// measurement never left it).
__global__ void on_copy_back(float* __restrict__ home, const float* __restrict__ cur, int R,
                             int64_t per_rep, const int64_t* __restrict__ n_meas_rep)
{
    const int rep = blockIdx.y;
    if (rep >= R || (n_meas_rep && n_meas_rep[rep] <= 0)) return;
    const float4* src = (const float4*)(cur + (int64_t)rep * per_rep);
    float4* dst = (float4*)(home + (int64_t)rep * per_rep);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < per_rep / 4;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// The workspace of cm_on_run: FP64 measurement partials (ON_CHUNK-site chunk
// partials), then for XY on square lattices a second spin buffer that the
// fused over-relaxation launch writes into (it runs out of place, alternating
// between the caller's buffer and this one).
struct OnWs {
    int64_t alt, total;
};
static OnWs on_ws_layout(int64_t N, int n_rep, int n, int d)
{
    const int64_t nchunk = (N + ON_CHUNK - 1) / ON_CHUNK;
    OnWs w;
    w.alt = ((int64_t)n_rep * nchunk * (1 + n + 2 * d) * 8 + 255) / 256 * 256;
    w.total = w.alt + (n == 2 && d == 2 ? (int64_t)n_rep * N * n * 4 : 0);
    return w;
}

// Renormalise every spin in FP64 (critmc/kernels/on.py:78-87).
template <int NC>
__global__ void on_normalize(float* __restrict__ spins, int64_t total_sites, int64_t N,
                             const int64_t* __restrict__ n_meas_rep, int64_t meas)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total_sites;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (n_meas_rep && meas >= n_meas_rep[t / N]) continue;
        const Vec<NC> s = unit_vec<NC>(load_vec<NC>(spins, t));
        store_vec<NC>(spins, t, s);
    }
}

// Partial FP64 sums of E, M[n], c_a, s_a over chunks of ON_CHUNK sites.
template <int NC, int D>
__global__ __launch_bounds__(256) void on_measure_partial(const float* __restrict__ spins, CbGeo g,
                                                          double* __restrict__ part, int nchunk,
                                                          const int64_t* __restrict__ n_meas_rep,
                                                          int64_t meas)
{
    constexpr int K = 1 + NC + 2 * D;
    __shared__ double red[K * 4];
    const int rep = blockIdx.y, chunk = blockIdx.x;
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;
    const float* S = spins + (int64_t)rep * g.N * NC;
    double v[K];
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = 0.0;
    const int64_t i0 = (int64_t)chunk * ON_CHUNK;
    const int64_t i1 = min(g.N, i0 + ON_CHUNK);
    const int64_t st[3] = {1, g.L0, (int64_t)g.L0 * g.L1};
    const int ext[3] = {g.L0, g.L1, g.L2};
    const FastDiv dv[2] = {make_fastdiv((uint32_t)g.L0), make_fastdiv((uint32_t)g.L1)};
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        int x[3] = {0, 0, 0};
        uint32_t rest = (uint32_t)i;
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            if (ax < 2) {
                const uint32_t q = fdiv(rest, dv[ax]);
                x[ax] = (int)(rest - q * (uint32_t)ext[ax]);
                rest = q;
            } else {
                x[ax] = (int)rest;
            }
        }
        const int par = (x[0] + x[1] + x[2]) & 1;
        const Vec<NC> s = load_vec<NC>(S, (int64_t)par * g.half + (i >> 1));
        const int64_t other = (int64_t)(1 - par) * g.half;
#pragma unroll
        for (int c = 0; c < NC; ++c) v[1 + c] += (double)s.v[c];
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            const int64_t j = i + (x[ax] + 1 == ext[ax] ? -(int64_t)(ext[ax] - 1) * st[ax] : st[ax]);
            const Vec<NC> t = load_vec<NC>(S, other + (j >> 1));
            double dot = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) dot += (double)s.v[c] * (double)t.v[c];
            v[0] -= dot;
            v[1 + NC + ax] += (double)s.v[0] * t.v[0] + (double)s.v[1] * t.v[1];
            v[1 + NC + D + ax] += (double)s.v[0] * t.v[1] - (double)s.v[1] * t.v[0];
        }
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const double t = wave_sum_f64(v[q]);
        if ((threadIdx.x & 63) == 0) red[q * 4 + w] = t;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double t = 0.0;
        for (int u = 0; u < (int)(blockDim.x >> 6); ++u) t += red[threadIdx.x * 4 + u];
        part[((int64_t)rep * nchunk + chunk) * K + threadIdx.x] = t;
    }
}

// The measurement terms of colour-0 site i, already renormalised as s: M of
// i and of i + e_0 (a bijection onto colour 1) and its 2d bonds, forward
// (i, i + e_a) and backward (i - e_a, i), oriented as the reference's forward
// bonds (every bond has exactly one colour-0 end).  v = (E, M[NC], c_a, s_a).
template <int NC, int D>
__device__ __forceinline__ void on_c0_bonds(const Vec<NC>& s, const float* __restrict__ S1,
                                            const uint32_t (&nbo)[2 * D], double (&v)[1 + NC + 2 * D])
{
    const Vec<NC> xr = load_vec<NC>(S1, nbo[0]);  // i + e_0
#pragma unroll
    for (int c = 0; c < NC; ++c) v[1 + c] += (double)s.v[c] + (double)xr.v[c];
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
#pragma unroll
        for (int dir = 0; dir < 2; ++dir) {
            const Vec<NC> t = load_vec<NC>(S1, nbo[2 * ax + dir]);
            const Vec<NC>& A = dir == 0 ? s : t;
            const Vec<NC>& B = dir == 0 ? t : s;
            double dot = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) dot += (double)A.v[c] * (double)B.v[c];
            v[0] -= dot;
            v[1 + NC + ax] += (double)A.v[0] * B.v[0] + (double)A.v[1] * B.v[1];
            v[1 + NC + D + ax] += (double)A.v[0] * B.v[1] - (double)A.v[1] * B.v[0];
        }
    }
}

// Renormalise-and-measure over the colour-0 sites, for measurements whose
// last half sweep (colour 1) already wrote its sites renormalised.  Every
// bond joins a colour-0 site i to a colour-1 site, so i's lane takes all 2d
// bonds of i (forward (i, i+e_a) and backward (i-e_a, i), oriented as the
// reference's forward bonds), renormalises i in place (no other lane reads
// it) and adds M of i and of i+e_0 (a bijection onto colour 1).  One
// workgroup per chunk of ON_CHUNK colour-0 sites, lane t taking sites
// t, t + 256, ... in turn, and on_measure_final adds the chunks in order: the
// colour-0 half sweep that carries a measurement (on_c0_sweep_measure) and
// the per-replica LDS form form the same partials, so every path gives the
// same bits.
template <int NC, int D>
__global__ __launch_bounds__(256) void on_measure_c0(float* __restrict__ spins, CbGeo g,
                                                     double* __restrict__ part, int nchunk,
                                                     const int64_t* __restrict__ n_meas_rep,
                                                     int64_t meas)
{
    constexpr int K = 1 + NC + 2 * D;
    constexpr int Z = 2 * D;
    __shared__ double red[K * 4];
    // consecutive chunks of a replica on one XCD, so the rows a chunk reads
    // beyond its own are in the same L2 as their neighbour chunk's
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x), chunk = (int)(lb - (uint32_t)rep * gridDim.x);
    if (n_meas_rep && meas >= n_meas_rep[rep]) return;
    // the colour-0 plane is renormalised in place, the colour-1 plane only read:
    // separate restrict pointers let the loads move ahead of the stores
    float* __restrict__ S0 = spins + (int64_t)rep * g.N * NC;
    const float* __restrict__ S1 = S0 + g.half * NC;
    double v[K];
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = 0.0;
    const int64_t k0 = (int64_t)chunk * ON_CHUNK;
    const int64_t k1 = min(g.half, k0 + ON_CHUNK);
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        uint32_t nbo[Z];
        cb_other<Z>(g, (uint32_t)k, 0, nbo);
        const Vec<NC> s = unit_vec<NC>(load_vec<NC>(S0, k));
        store_vec<NC>(S0, k, s);
        on_c0_bonds<NC, D>(s, S1, nbo, v);
    }
    on_block_partial<K>(v, red, part + ((int64_t)rep * nchunk + chunk) * K);
}

// The colour-0 half sweep that carries the previous measurement (a
// measurement ending with a colour-1 half sweep, the next starting with this
// one): per chunk of ON_CHUNK sites, lane t takes sites t, t + 256, ... as
// on_measure_c0 does, renormalises each and adds its bond and magnetisation
// terms to the chunk partial of measurement `pmeas` (the same terms in the
// same order, so the same bits), then updates it with on_site.  The reference
// renormalises at the measurement and starts the next sweep from those spins
// (critmc/kernels/on.py:404-413), so the chain is the same.  A replica whose
// cap ends before `a.meas` is only measured; one capped before `pmeas` is
// left alone.
template <int NC, int Z, int CODE>
__global__ __launch_bounds__(256) void on_c0_sweep_measure(OnArgs a, double* __restrict__ part,
                                                           int nchunk, int64_t pmeas)
{
    constexpr int D = Z / 2, K = 1 + NC + Z;
    __shared__ double red[K * 4];
    __shared__ int ired[4];
    const uint32_t lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int rep = (int)(lb / gridDim.x), chunk = (int)(lb - (uint32_t)rep * gridDim.x);
    const bool upd = !(a.n_meas_rep && a.meas >= a.n_meas_rep[rep]);
    const bool msr = !(a.n_meas_rep && pmeas >= a.n_meas_rep[rep]);
    if (!upd && !msr) return;  // whole block
    float* __restrict__ S0 = a.spins + (int64_t)rep * a.g.N * NC;
    const float* __restrict__ S1 = S0 + a.g.half * NC;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const float beta = a.betas[rep], step = a.steps[rep];
    double v[K];
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = 0.0;
    int accepted = 0;
    const int64_t c0 = (int64_t)chunk * ON_CHUNK;
    const int64_t c1 = min(a.g.half, c0 + ON_CHUNK);
    for (int64_t k = c0 + threadIdx.x; k < c1; k += blockDim.x) {
        uint32_t nbo[Z];
        cb_other<Z>(a.g, (uint32_t)k, 0, nbo);
        Vec<NC> s = load_vec<NC>(S0, k);
        bool changed = false;
        if (msr) {
            s = unit_vec<NC>(s);
            on_c0_bonds<NC, D>(s, S1, nbo, v);
            changed = true;
        }
        if (upd) {
            int ac = 0;  // on_site sets it to 1 on an accepted proposal
            changed |= on_site<NC, Z, CODE>(s, S1, nbo, (uint32_t)k, 0, sweep, k0, k1, sid, beta, step, ac);
            accepted += ac;
        }
        if (changed) store_vec<NC>(S0, k, s);
    }
    if (msr) on_block_partial<K>(v, red, part + ((int64_t)rep * nchunk + chunk) * K);
    if (CODE == CODE_METROPOLIS && upd) {  // one atomic per workgroup
        const int t = wave_sum(accepted);
        if ((threadIdx.x & 63) == 0) ired[threadIdx.x >> 6] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int tot = ired[0] + ired[1] + ired[2] + ired[3];
            if (tot) atomicAdd(&a.acc[(int64_t)rep * a.n_meas + a.meas], tot);
        }
    }
}

// Per-replica form for lattices whose spins fit in LDS (N n 4 B <= ON_BATCHED_LDS):
// one workgroup per replica runs every sweep and measurement of the call with
// the replica in LDS, in the colour-split layout, and writes it back once.
// Each site update is on_site, and each measurement repeats on_measure_c0's
// per-site work, chunks and reductions and then on_measure_final's ordered
// sum, so spins and series equal the per-half-sweep launches' bit for bit.
// Local updates only (Metropolis, heat bath, over-relaxation); as there, the
// last half sweep of a measurement stores its sites renormalised.  HB: the
// schedule has heat-bath rows (a separate build, as the heat bath's registers
// would otherwise limit every schedule's occupancy).  BLK lanes per workgroup:
// 256, or 1024 for replicas large enough to give each lane several sites
// (local updates without heat bath); the measurement then runs in groups of
// 256 lanes, group j taking chunk j of every NG = BLK / 256 consecutive
// chunks exactly as one on_measure_c0 workgroup takes it, and the chunk
// partials are added in chunk order, so the bits do not depend on BLK.
// Up to 60 KB of spins any number of replicas take this form (two or more
// workgroups per CU); up to 158 KB (XY 128^2 at 128 KB, one workgroup per CU)
// only as many replicas as the GPU has CUs.  More such replicas would hold
// every CU for several rounds of a whole call and starve work on other streams
// (the all-queued serving stream lost about 10 % that way); they take the
// half-sweep launches instead.
constexpr int64_t ON_BATCHED_LDS_ANY = 60 * 1024;
constexpr int64_t ON_BATCHED_LDS = 158 * 1024;

// CUs of the current device (cached per device).
static int on_cu_count()
{
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if (cached[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            cached[dev] = v;
    }
    return cached[dev];
}

template <int NC, int D, bool HB, int BLK>
__global__ __launch_bounds__(BLK) void on_batched(OnArgs a, Sched sc, double* __restrict__ e_out,
                                                  double* __restrict__ m_out,
                                                  double* __restrict__ sc_out,
                                                  double* __restrict__ ss_out)
{
    constexpr int Z = 2 * D, K = 1 + NC + 2 * D;
    constexpr int NG = BLK / 256;  // measurement groups of 256 lanes
    static_assert(BLK % 256 == 0, "the measurement takes chunks in groups of 256 lanes");
    extern __shared__ float ls[];  // [2 colours][half][NC]
    __shared__ double red[K * NG * 4];
    __shared__ int ired[BLK / 64];
    const CbGeo& g = a.g;
    const int64_t half = g.half;
    const int w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    float* L0 = ls;
    float* L1 = ls + half * NC;
    for (int rep = blockIdx.x; rep < a.R; rep += gridDim.x) {
        const int64_t nmr = a.n_meas_rep ? min(a.n_meas, a.n_meas_rep[rep]) : a.n_meas;
        if (nmr <= 0) continue;  // the whole block: replica capped
        float* S = a.spins + (int64_t)rep * g.N * NC;
        for (int64_t t = threadIdx.x; t < g.N * NC; t += blockDim.x) ls[t] = S[t];
        __syncthreads();
        const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
        const uint64_t sw0 = a.sweeps[rep];
        const float beta = a.betas[rep], step = a.steps[rep];
        uint64_t sw = 0;  // sweeps of this call before the current one
        for (int64_t meas = 0; meas < nmr; ++meas) {
            int acc = 0;
            for (int r = 0; r < sc.n; ++r) {
                const int code = sc.code[r];
                for (int rp = 0; rp < sc.count[r]; ++rp, ++sw) {
                    for (int c = 0; c < 2; ++c) {
                        const bool norm = r == sc.n - 1 && rp == sc.count[r] - 1 && c == 1;
                        float* Pc = c ? L1 : L0;
                        const float* Po = c ? L0 : L1;
                        for (int64_t k = threadIdx.x; k < half; k += blockDim.x) {
                            uint32_t nbo[Z];
                            cb_other<Z>(g, (uint32_t)k, c, nbo);
                            Vec<NC> s = load_vec<NC>(Pc, k);
                            int ac = 0;
                            bool changed;
                            if (code == CODE_METROPOLIS)
                                changed = on_site<NC, Z, CODE_METROPOLIS>(s, Po, nbo, (uint32_t)k, c, sw0 + sw,
                                                                          k0, k1, sid, beta, step, ac);
                            else if (HB && code == CODE_HEATBATH)
                                changed = on_site<NC, Z, CODE_HEATBATH>(s, Po, nbo, (uint32_t)k, c, sw0 + sw,
                                                                        k0, k1, sid, beta, step, ac);
                            else
                                changed = on_site<NC, Z, CODE_OVERRELAX>(s, Po, nbo, (uint32_t)k, c, sw0 + sw,
                                                                         k0, k1, sid, beta, step, ac);
                            acc += ac;
                            if (norm) {
                                s = unit_vec<NC>(s);
                                store_vec<NC>(Pc, k, s);
                            } else if (changed) {
                                store_vec<NC>(Pc, k, s);
                            }
                        }
                        __syncthreads();
                    }
                }
            }
            // the measurement: colour 0 renormalised in place, bonds summed per
            // chunk of ON_CHUNK sites by a group of 256 lanes, as on_measure_c0
            double total = 0.0;  // thread q < K: quantity q, chunk partials added in order
            const int nchunk = (int)((half + ON_CHUNK - 1) / ON_CHUNK);
            const int grp = (int)(threadIdx.x >> 8), gt = (int)(threadIdx.x & 255), gw = (gt >> 6);
            for (int cb = 0; cb < nchunk; cb += NG) {
                double v[K];
#pragma unroll
                for (int q = 0; q < K; ++q) v[q] = 0.0;
                const int64_t c0 = (int64_t)(cb + grp) * ON_CHUNK, c1 = min(half, c0 + ON_CHUNK);
                for (int64_t k = c0 + gt; k < c1; k += 256) {
                    uint32_t nbo[Z];
                    cb_other<Z>(g, (uint32_t)k, 0, nbo);
                    const Vec<NC> s = unit_vec<NC>(load_vec<NC>(L0, k));
                    store_vec<NC>(L0, k, s);
                    on_c0_bonds<NC, D>(s, L1, nbo, v);
                }
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const double t = wave_sum_f64(v[q]);
                    if ((threadIdx.x & 63) == 0) red[(q * NG + grp) * 4 + gw] = t;
                }
                __syncthreads();
                if (threadIdx.x < K) {
                    for (int j = 0; j < NG && cb + j < nchunk; ++j) {
                        double t = 0.0;
                        for (int u = 0; u < 4; ++u) t += red[((int)threadIdx.x * NG + j) * 4 + u];
                        total += t;
                    }
                }
                __syncthreads();
            }
            const int ta = wave_sum(acc);
            if ((threadIdx.x & 63) == 0) ired[w] = ta;
            const int64_t o = (int64_t)rep * a.n_meas + meas;
            if (threadIdx.x < K) {
                const int q = threadIdx.x;
                if (q == 0) e_out[o] = total;
                else if (q < 1 + NC) m_out[o * NC + (q - 1)] = total;
                else if (q < 1 + NC + D) sc_out[o * D + (q - 1 - NC)] = total;
                else ss_out[o * D + (q - 1 - NC - D)] = total;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int tot = 0;
                for (int u = 0; u < nw; ++u) tot += ired[u];
                a.acc[o] = tot;
            }
        }
        for (int64_t t = threadIdx.x; t < g.N * NC; t += blockDim.x) S[t] = ls[t];
        __syncthreads();
    }
}

// One wave per (replica, quantity): launch with (R K + 3) / 4 blocks of 256.
template <int NC, int D>
__global__ __launch_bounds__(256) void on_measure_final(const double* __restrict__ part, int nb, int R,
                                                        int64_t n_meas, int64_t meas, double* __restrict__ e_out,
                                                        double* __restrict__ m_out, double* __restrict__ sc_out,
                                                        double* __restrict__ ss_out,
                                                        const int64_t* __restrict__ n_meas_rep)
{
    on_final_pair<NC, D>(part, nb, R, n_meas, meas, e_out, m_out, sc_out, ss_out, n_meas_rep,
                         (int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
}

__global__ void fill_i32_on(int32_t* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void on_advance_sweeps(uint64_t* __restrict__ sweeps, int R, uint64_t per_meas,
                                  int64_t n_meas, const int64_t* __restrict__ n_meas_rep)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < R) {
        const int64_t n = n_meas_rep ? min(n_meas, n_meas_rep[t]) : n_meas;
        sweeps[t] += per_meas * (uint64_t)n;
    }
}

struct OnClArgs {
    float* spins;
    const float* betas;
    const uint32_t* keys;
    const uint32_t* sids;
    const uint64_t* sweeps;
    uint64_t sweep_add;
    int32_t* parent;  // [R][N]
    int32_t* size;    // [R][N]
    int32_t* queue;   // [R][N]
    int32_t* mark;    // [R][N]
    double* proj;     // [R][N]
    int64_t* stats;   // [R][2]: sum |C|^2, clusters of the last Swendsen-Wang update
    int64_t* wsize;   // [R]: Wolff sizes of the current measurement
    ClGeo g;
    int R;
};

template <int NC>
__device__ __forceinline__ void on_axis(uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep,
                                        double (&r)[NC])
{
    const u32x4 b = philox4x32_10(u32x4{0u, (uint32_t)sweep, ctr_z(sweep, 3, TAG_CLUSTER), sid}, k0, k1);
    if constexpr (NC == 2) {
        const double phi = 2.0 * M_PI * ((double)b.x * 0x1p-32);
        r[0] = cos(phi);
        r[1] = sin(phi);
    } else {
        const double z = 2.0 * ((double)b.x * 0x1p-32) - 1.0;
        const double az = 2.0 * M_PI * ((double)b.y * 0x1p-32);
        const double w = sqrt(fmax(0.0, 1.0 - z * z));
        r[0] = w * cos(az);
        r[1] = w * sin(az);
        r[2] = z;
    }
}

// Position of site i in the colour-split layout.
__device__ __forceinline__ uint32_t on_split(const ClGeo& g, uint32_t i)
{
    uint32_t x[3];
    cl_coords(g, i, x);
    return ((x[0] + x[1] + x[2]) & 1u) * (g.N >> 1) + (i >> 1);
}

template <int NC>
__device__ __forceinline__ double on_proj(const float* S, uint32_t p_i, const double (&r)[NC])
{
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) p += (double)S[(int64_t)NC * p_i + c] * r[c];
    return p;
}

template <int NC>
__device__ __forceinline__ void on_reflect(float* S, uint32_t p_i, double p, const double (&r)[NC])
{
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float* q = S + (int64_t)NC * p_i + c;
        *q = (float)((double)*q - 2.0 * p * r[c]);
    }
}

__device__ __forceinline__ bool on_bond_active(double x, double beta, uint32_t w)
{
    return x > 0.0 && (double)w * 0x1p-32 < 1.0 - det_exp(-2.0 * beta * x);
}

template <int NC>
__global__ __launch_bounds__(256) void on_sw_prep(OnClArgs a)
{
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    const float* S = a.spins + (int64_t)rep * N * NC;
    int32_t* P = a.parent + (int64_t)rep * N;
    int32_t* Z = a.size + (int64_t)rep * N;
    double* Pj = a.proj + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    double r[NC];
    on_axis<NC>(a.keys[2 * rep], a.keys[2 * rep + 1], a.sids[rep], sweep, r);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        P[i] = (int32_t)i;
        Z[i] = 0;
        Pj[i] = on_proj<NC>(S, on_split(a.g, i), r);
    }
    if (blockIdx.x == 0 && threadIdx.x < 2) a.stats[2 * rep + threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void on_sw_union(OnClArgs a)
{
    const int rep = blockIdx.y;
    const ClGeo& g = a.g;
    const uint32_t N = g.N, nb = (uint32_t)g.d * N;
    const double* Pj = a.proj + (int64_t)rep * N;
    int32_t* P = a.parent + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const double beta = (double)a.betas[rep];
    const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
    for (uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x; 4 * blk < nb;
         blk += gridDim.x * blockDim.x) {
        const u32x4 r = philox4x32_10(u32x4{blk, (uint32_t)sweep, zc, sid}, k0, k1);
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t b = 4 * blk + t;
            if (b >= nb) break;
            const int ax = (int)(b >= N) + (int)(b >= 2 * N);
            const uint32_t i = b - (uint32_t)ax * N;
            uint32_t x[3];
            cl_coords(g, i, x);
            const uint32_t j = cl_nbr(g, i, x, ax, true);
            if (on_bond_active(Pj[i] * Pj[j], beta, w[t])) cl_unite(P, i, j);
        }
    }
}

__global__ __launch_bounds__(256) void on_sw_flatten(OnClArgs a)
{
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    int32_t* P = a.parent + (int64_t)rep * N;
    int32_t* Z = a.size + (int64_t)rep * N;
    // near the critical point one cluster holds many of the sites: up to three
    // times, the lanes of a wave that share its first remaining lane's root
    // add their count in one atomic (integer sums, so the order is free)
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < N; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;  // i0 is uniform, so the wave stays together
        const bool site = i < N;
        uint32_t r = 0;
        if (site) {
            r = cl_find_ro(P, i);
            __hip_atomic_store(P + i, (int32_t)r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        bool left = site;
#pragma unroll
        for (int pass = 0; pass < 3; ++pass) {
            const uint64_t rem = __ballot(left);
            if (rem == 0) break;
            const int first = __ffsll((long long)rem) - 1;
            const uint32_t r0 = (uint32_t)__shfl((int)r, first, 64);
            const bool mine = left && r == r0;
            const uint64_t grp = __ballot(mine);
            if (mine && (int)(threadIdx.x & 63) == first) atomicAdd(Z + r0, __popcll(grp));
            left = left && !mine;
        }
        if (left) atomicAdd(Z + r, 1);
    }
}

template <int NC>
__global__ __launch_bounds__(256) void on_sw_flip(OnClArgs a)
{
    __shared__ long long red[2][4];
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    float* S = a.spins + (int64_t)rep * N * NC;
    const int32_t* P = a.parent + (int64_t)rep * N;
    const int32_t* Z = a.size + (int64_t)rep * N;
    const double* Pj = a.proj + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    double r[NC];
    on_axis<NC>(k0, k1, sid, sweep, r);
    long long c2 = 0, ncl = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint32_t rt = (uint32_t)P[i];
        if (rt == i) {
            const long long z = Z[i];
            c2 += z * z;
            ncl += 1;
        }
        if ((cl_word(k0, k1, sid, sweep, 1, rt) >> 31) == 0) on_reflect<NC>(S, on_split(a.g, i), Pj[i], r);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c2 += __shfl_xor(c2, off, 64);
        ncl += __shfl_xor(ncl, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = c2;
        red[1][threadIdx.x >> 6] = ncl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t0 = 0, t1 = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            t0 += red[0][w];
            t1 += red[1][w];
        }
        if (t0) atomicAdd((unsigned long long*)&a.stats[2 * rep], (unsigned long long)t0);
        if (t1) atomicAdd((unsigned long long*)&a.stats[2 * rep + 1], (unsigned long long)t1);
    }
}

template <int NC>
__global__ __launch_bounds__(256) void on_wolff(OnClArgs a, int count)
{
    __shared__ uint32_t s_head, s_tail;
    const int rep = blockIdx.x;
    const ClGeo& g = a.g;
    const uint32_t N = g.N;
    float* S = a.spins + (int64_t)rep * N * NC;
    int32_t* Q = a.queue + (int64_t)rep * N;
    int32_t* Mk = a.mark + (int64_t)rep * N;
    double* Pj = a.proj + (int64_t)rep * N;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const double beta = (double)a.betas[rep];
    const int z = 2 * g.d;
    long long total = 0;
    for (int f = 0; f < count; ++f) {
        const uint64_t sweep = a.sweeps[rep] + a.sweep_add + (uint64_t)f;
        double r[NC];
        on_axis<NC>(k0, k1, sid, sweep, r);
        if (threadIdx.x == 0) {
            const uint32_t x = cl_word(k0, k1, sid, sweep, 2, 0);
            const uint32_t seed = (uint32_t)(((uint64_t)x * N) >> 32);
            const uint32_t ps = on_split(g, seed);
            const double p0 = on_proj<NC>(S, ps, r);
            Pj[seed] = p0;
            on_reflect<NC>(S, ps, p0, r);
            Mk[seed] = f;
            Q[0] = (int32_t)seed;
            s_head = 0;
            s_tail = 1;
        }
        __syncthreads();
        const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
        while (true) {
            const uint32_t lo = s_head, hi = s_tail;
            // every lane has this level's bounds before any lane appends to the queue
            __syncthreads();
            if (lo == hi) break;
            for (uint32_t q = lo + threadIdx.x; q < hi; q += blockDim.x) {
                const uint32_t k = (uint32_t)Q[q];
                const double pk = Pj[k];
                uint32_t x[3];
                cl_coords(g, k, x);
                for (int c = 0; c < z; ++c) {
                    const int ax = c >> 1;
                    const uint32_t j = cl_nbr(g, k, x, ax, (c & 1) == 0);
                    if (__hip_atomic_load(Mk + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == f)
                        continue;
                    const uint32_t sj = on_split(g, j);
                    const double pj = on_proj<NC>(S, sj, r);
                    const uint32_t b = (uint32_t)ax * N + ((c & 1) ? j : k);
                    const u32x4 rn = philox4x32_10(u32x4{b >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
                    const uint32_t w = b & 3;
                    const uint32_t u = w == 0 ? rn.x : (w == 1 ? rn.y : (w == 2 ? rn.z : rn.w));
                    if (!on_bond_active(pk * pj, beta, u)) continue;
                    if (atomicExch(Mk + j, f) == f) continue;  // another lane added it
                    Pj[j] = pj;
                    on_reflect<NC>(S, sj, pj, r);
                    Q[atomicAdd(&s_tail, 1u)] = (int32_t)j;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) s_head = hi;
            __syncthreads();
        }
        total += s_tail;
        __syncthreads();
    }
    if (threadIdx.x == 0) a.wsize[rep] += total;
}

__global__ void on_cl_fill(int32_t* p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void on_cl_zero64(int64_t* p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

__global__ void on_cl_record(int R, int64_t n_meas, int64_t meas, const int32_t* acc,
                             const int64_t* wsize, const int64_t* stats, int have_sw,
                             int64_t* aux)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= R) return;
    const int64_t o = (int64_t)rep * n_meas + meas;
    aux[4 * o + 0] = acc[o];
    aux[4 * o + 1] = wsize ? wsize[rep] : 0;  // null without cluster rows
    aux[4 * o + 2] = have_sw && stats ? stats[2 * rep] : 0;
    aux[4 * o + 3] = have_sw && stats ? stats[2 * rep + 1] : 0;
}

inline ClGeo on_clgeo(int32_t ndim, const int32_t* dims)
{
    ClGeo c{};
    c.d = ndim;
    int64_t N = 1;
    for (int k = 0; k < 3; ++k) {
        c.L[k] = k < ndim ? (uint32_t)dims[k] : 1u;
        N *= c.L[k];
    }
    c.N = (uint32_t)N;
    c.div0 = make_fastdiv(c.L[0]);
    c.div1 = make_fastdiv(c.L[1]);
    return c;
}

// Site order [R][N][n] <-> colour-split [R][2][N/2][n]; direction 0: to split.
__global__ void on_convert_kernel(const float* __restrict__ src, float* __restrict__ dst, ClGeo g,
                                  int NC, int direction)
{
    const int rep = blockIdx.y;
    const int64_t N = g.N;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += gridDim.x * blockDim.x) {
        const int64_t a = ((int64_t)rep * N + i) * NC;
        const int64_t b = ((int64_t)rep * N + on_split(g, i)) * NC;
        for (int c = 0; c < NC; ++c) {
            if (direction == 0) dst[b + c] = src[a + c];
            else dst[a + c] = src[b + c];
        }
    }
}

struct OnClLayout {
    int64_t parent, size, queue, mark, proj, stats, wsize, total;
};

inline OnClLayout on_cl_layout(int64_t N, int32_t n_rep)
{
    auto up = [](int64_t x) { return (x + 255) / 256 * 256; };
    OnClLayout l{};
    int64_t o = 0;
    l.parent = o; o += up(4 * N * n_rep);
    l.size = o;   o += up(4 * N * n_rep);
    l.queue = o;  o += up(4 * N * n_rep);
    l.mark = o;   o += up(4 * N * n_rep);
    l.proj = o;   o += up(8 * N * n_rep);
    l.stats = o;  o += up(16 * (int64_t)n_rep);
    l.wsize = o;  o += up(8 * (int64_t)n_rep);
    l.total = o;
    return l;
}

template <int NC>
static void on_cluster_row(OnClArgs a, int64_t code, int64_t count, hipStream_t st)
{
    const int64_t N = a.g.N;
    const unsigned site_chunks = (unsigned)std::min<int64_t>((N + 255) / 256, 1024);
    const unsigned bond_chunks = (unsigned)std::min<int64_t>((a.g.d * N / 4 + 256) / 256, 1024);
    const unsigned flip_chunks = (unsigned)std::min<int64_t>((N + 4095) / 4096, 1024);
    if (code == CODE_SWENDSEN_WANG) {
        const uint64_t base = a.sweep_add;
        for (int64_t k = 0; k < count; ++k) {
            a.sweep_add = base + (uint64_t)k;
            hipLaunchKernelGGL((on_sw_prep<NC>), dim3(site_chunks, a.R), dim3(256), 0, st, a);
            hipLaunchKernelGGL(on_sw_union, dim3(bond_chunks, a.R), dim3(256), 0, st, a);
            hipLaunchKernelGGL(on_sw_flatten, dim3(site_chunks, a.R), dim3(256), 0, st, a);
            // 16 sites per lane: fewer workgroup-end atomics on the statistics
            hipLaunchKernelGGL((on_sw_flip<NC>), dim3(flip_chunks, a.R), dim3(256), 0, st, a);
        }
    } else {
        const int64_t nm = N * a.R;
        hipLaunchKernelGGL(on_cl_fill, dim3((unsigned)std::min<int64_t>((nm + 255) / 256, 8192)),
                           dim3(256), 0, st, a.mark, nm, -1);
        hipLaunchKernelGGL((on_wolff<NC>), dim3(a.R), dim3(256), 0, st, a, (int)count);
    }
}

__global__ void on_init_kernel(float* __restrict__ spins, ClGeo g, int NC,
                               const uint32_t* __restrict__ keys,
                               const uint32_t* __restrict__ sids)
{
    const int64_t N = g.N;
    const int rep = blockIdx.y;
    const uint32_t k0 = keys[2 * rep], k1 = keys[2 * rep + 1], sid = sids[rep];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        const u32x4 r = draw(k0, k1, sid, 0, 0, TAG_INIT, (uint32_t)i);
        float* p = spins + ((int64_t)rep * N + on_split(g, (uint32_t)i)) * NC;
        if (NC == 2) {
            // angle uniform in [0, 2 pi): theta = 2 pi x / 2^32
            float sn, cs;
            sincospif(2.0f * ((float)r.x * 0x1p-32f), &sn, &cs);
            p[0] = cs;
            p[1] = sn;
        } else {
            // isotropic: z uniform in [-1, 1), azimuth uniform
            const float z = 2.0f * ((float)r.x * 0x1p-32f) - 1.0f;
            const float rho = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            float sn, cs;
            sincospif(2.0f * ((float)r.y * 0x1p-32f), &sn, &cs);
            p[0] = rho * cs;
            p[1] = rho * sn;
            p[2] = z;
        }
    }
}

// A half sweep over all replicas, one colour site per lane; with fin.part set
// one more row of workgroups takes that earlier measurement's ordered sums.
template <int NC, int Z, bool FIN>
static void launch_half(int code, const OnArgs& a, const OnFinal& fin, hipStream_t st)
{
    const dim3 grid((unsigned)((a.g.half + 255) / 256), a.R + (FIN ? 1 : 0));
    if (code == CODE_OVERRELAX)
        hipLaunchKernelGGL((on_half_sweep<NC, Z, CODE_OVERRELAX, FIN>), grid, dim3(256), 0, st, a, fin);
    else if (code == CODE_HEATBATH)
        hipLaunchKernelGGL((on_half_sweep<NC, Z, CODE_HEATBATH, FIN>), grid, dim3(256), 0, st, a, fin);
    else
        hipLaunchKernelGGL((on_half_sweep<NC, Z, CODE_METROPOLIS, FIN>), grid, dim3(256), 0, st, a, fin);
}

template <int NC>
static void launch_half_z(int code, const OnArgs& a, hipStream_t st, const OnFinal& fin = OnFinal{})
{
    if (fin.part) {
        if (a.g.d == 1) launch_half<NC, 2, true>(code, a, fin, st);
        else if (a.g.d == 2) launch_half<NC, 4, true>(code, a, fin, st);
        else launch_half<NC, 6, true>(code, a, fin, st);
        return;
    }
    if (a.g.d == 1) launch_half<NC, 2, false>(code, a, fin, st);
    else if (a.g.d == 2) launch_half<NC, 4, false>(code, a, fin, st);
    else launch_half<NC, 6, false>(code, a, fin, st);
}

// The colour-0 half sweep that carries measurement `pmeas` (on_c0_sweep_measure).
template <int NC, int Z>
static void launch_c0_measure_z(int code, const OnArgs& a, double* part, int nchunk, int64_t pmeas,
                                hipStream_t st)
{
    const dim3 grid((unsigned)nchunk, a.R);
    if (code == CODE_OVERRELAX)
        hipLaunchKernelGGL((on_c0_sweep_measure<NC, Z, CODE_OVERRELAX>), grid, dim3(256), 0, st, a, part, nchunk, pmeas);
    else if (code == CODE_HEATBATH)
        hipLaunchKernelGGL((on_c0_sweep_measure<NC, Z, CODE_HEATBATH>), grid, dim3(256), 0, st, a, part, nchunk, pmeas);
    else
        hipLaunchKernelGGL((on_c0_sweep_measure<NC, Z, CODE_METROPOLIS>), grid, dim3(256), 0, st, a, part, nchunk, pmeas);
}

template <int NC>
static void launch_c0_measure(int code, const OnArgs& a, double* part, int nchunk, int64_t pmeas,
                              hipStream_t st)
{
    if (a.g.d == 1) launch_c0_measure_z<NC, 2>(code, a, part, nchunk, pmeas, st);
    else if (a.g.d == 2) launch_c0_measure_z<NC, 4>(code, a, part, nchunk, pmeas, st);
    else launch_c0_measure_z<NC, 6>(code, a, part, nchunk, pmeas, st);
}

template <int NC>
static void launch_measure(const float* spins, const CbGeo& g, int R, double* part, int nchunk,
                           int64_t n_meas, int64_t meas, double* e_out, double* m_out,
                           double* sc_out, double* ss_out, const int64_t* nr, hipStream_t st)
{
    const dim3 grid(nchunk, R);
    if (g.d == 1) {
        hipLaunchKernelGGL((on_measure_partial<NC, 1>), grid, dim3(256), 0, st, spins, g, part, nchunk, nr, meas);
        hipLaunchKernelGGL((on_measure_final<NC, 1>), dim3((R * (1 + NC + 2) + 3) / 4), dim3(256), 0, st, part, nchunk, R, n_meas, meas, e_out, m_out, sc_out, ss_out, nr);
    } else if (g.d == 2) {
        hipLaunchKernelGGL((on_measure_partial<NC, 2>), grid, dim3(256), 0, st, spins, g, part, nchunk, nr, meas);
        hipLaunchKernelGGL((on_measure_final<NC, 2>), dim3((R * (1 + NC + 4) + 3) / 4), dim3(256), 0, st, part, nchunk, R, n_meas, meas, e_out, m_out, sc_out, ss_out, nr);
    } else {
        hipLaunchKernelGGL((on_measure_partial<NC, 3>), grid, dim3(256), 0, st, spins, g, part, nchunk, nr, meas);
        hipLaunchKernelGGL((on_measure_final<NC, 3>), dim3((R * (1 + NC + 6) + 3) / 4), dim3(256), 0, st, part, nchunk, R, n_meas, meas, e_out, m_out, sc_out, ss_out, nr);
    }
}

template <int NC>
static void launch_measure_c0(float* spins, const CbGeo& g, int R, double* part, int64_t n_meas,
                              int64_t meas, double* e_out, double* m_out, double* sc_out,
                              double* ss_out, const int64_t* nr, hipStream_t st)
{
    const int nc0 = (int)((g.half + ON_CHUNK - 1) / ON_CHUNK);
    const dim3 grid(nc0, R);
    if (g.d == 1) {
        hipLaunchKernelGGL((on_measure_c0<NC, 1>), grid, dim3(256), 0, st, spins, g, part, nc0, nr, meas);
        hipLaunchKernelGGL((on_measure_final<NC, 1>), dim3((R * (1 + NC + 2) + 3) / 4), dim3(256), 0, st, part, nc0, R, n_meas, meas, e_out, m_out, sc_out, ss_out, nr);
    } else if (g.d == 2) {
        hipLaunchKernelGGL((on_measure_c0<NC, 2>), grid, dim3(256), 0, st, spins, g, part, nc0, nr, meas);
        hipLaunchKernelGGL((on_measure_final<NC, 2>), dim3((R * (1 + NC + 4) + 3) / 4), dim3(256), 0, st, part, nc0, R, n_meas, meas, e_out, m_out, sc_out, ss_out, nr);
    } else {
        hipLaunchKernelGGL((on_measure_c0<NC, 3>), grid, dim3(256), 0, st, spins, g, part, nc0, nr, meas);
        hipLaunchKernelGGL((on_measure_final<NC, 3>), dim3((R * (1 + NC + 6) + 3) / 4), dim3(256), 0, st, part, nc0, R, n_meas, meas, e_out, m_out, sc_out, ss_out, nr);
    }
}

template <int NC, int D, bool HB, int BLK>
static hipError_t launch_batched_one(dim3 grid, size_t lds, hipStream_t st, const OnArgs& a,
                                     const Sched& sc, double* e_out, double* m_out, double* sc_out,
                                     double* ss_out)
{
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void*)on_batched<NC, D, HB, BLK>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((on_batched<NC, D, HB, BLK>), grid, dim3(BLK), lds, st, a, sc, e_out, m_out,
                       sc_out, ss_out);
    return hipGetLastError();
}

// 1024 lanes once a colour has 4096 sites (four or more per lane), local
// updates without heat bath (whose registers need the 256-lane build).
template <int NC, int D, bool HB>
static hipError_t launch_batched_blk(dim3 grid, size_t lds, hipStream_t st, const OnArgs& a,
                                     const Sched& sc, double* e_out, double* m_out, double* sc_out,
                                     double* ss_out)
{
    if (!HB && a.g.half >= 4096)
        return launch_batched_one<NC, D, HB, 1024>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
    return launch_batched_one<NC, D, HB, 256>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
}

template <bool HB>
static hipError_t launch_batched_on(int n, int d, dim3 grid, size_t lds, hipStream_t st,
                                    const OnArgs& a, const Sched& sc, double* e_out, double* m_out,
                                    double* sc_out, double* ss_out)
{
    if (n == 2) {
        if (d == 1) return launch_batched_blk<2, 1, HB>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
        if (d == 2) return launch_batched_blk<2, 2, HB>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
        return launch_batched_blk<2, 3, HB>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
    }
    if (d == 1) return launch_batched_blk<3, 1, HB>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
    if (d == 2) return launch_batched_blk<3, 2, HB>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
    return launch_batched_blk<3, 3, HB>(grid, lds, st, a, sc, e_out, m_out, sc_out, ss_out);
}

}  // namespace cm

using namespace cm;

extern "C" {

int64_t cm_on_workspace_bytes(int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t n)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    return on_ws_layout(N, n_rep, n, ndim).total;
}

int64_t cm_on_cluster_workspace_bytes(int32_t ndim, const int32_t* dims, int32_t n_rep)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    return on_cl_layout(N, n_rep).total;
}

int cm_on_init(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t n,
               const uint32_t* keys, const uint32_t* stream_ids, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(n == 2 || n == 3, "the GPU O(n) kernels support n = 2 (XY) and n = 3 (Heisenberg)");
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    const int blocks = (int)std::min<int64_t>((g.N + 255) / 256, 1024);
    hipLaunchKernelGGL(on_init_kernel, dim3(blocks, n_rep), dim3(256), 0, (hipStream_t)stream,
                       (float*)spins, on_clgeo(ndim, dims), n, keys, stream_ids);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_on_convert(const void* src, void* dst, int32_t ndim, const int32_t* dims, int32_t n_rep,
                  int32_t n, int32_t direction, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(n == 2 || n == 3, "n must be 2 or 3");
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    CM_REQUIRE(direction == 0 || direction == 1, "direction must be 0 (to split) or 1 (to sites)");
    CM_REQUIRE(src != dst, "conversion is out of place");
    const int blocks = (int)std::min<int64_t>((g.N + 255) / 256, 1024);
    hipLaunchKernelGGL(on_convert_kernel, dim3(blocks, n_rep), dim3(256), 0, (hipStream_t)stream,
                       (const float*)src, (float*)dst, on_clgeo(ndim, dims), n, direction);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_on_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t n,
              const float* betas_dev, const float* steps_dev, const uint32_t* keys,
              const uint32_t* stream_ids, uint64_t* sweeps, const int64_t* schedule,
              int32_t n_rows, int64_t n_meas, int32_t* acc_out, double* e_out, double* m_out,
              double* sc_out, double* ss_out, void* workspace, int64_t workspace_bytes,
              const int64_t* n_meas_rep, int64_t* aux_out, void* cluster_ws,
              int64_t cluster_ws_bytes, void* stream)
{
    CbGeo g;
    CM_REQUIRE(cb_geo_from(ndim, dims, &g) == 0, "checkerboard sweeps need 1..3 even extents >= 2");
    CM_REQUIRE(n == 2 || n == 3, "the GPU O(n) kernels support n = 2 (XY) and n = 3 (Heisenberg)");
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    CM_REQUIRE(n_meas >= 0, "n_meas must not be negative");
    CM_REQUIRE(n_rows >= 1 && n_rows <= 8, "schedule must have 1..8 rows");
    CM_REQUIRE(g.half <= 0xFFFFFFFFll / 2, "lattice too large");
    CM_REQUIRE((g.half + 255) / 256 * n_rep < (1ll << 32), "too many sites x replicas per launch");
    CM_REQUIRE(workspace_bytes >= cm_on_workspace_bytes(ndim, dims, n_rep, n), "workspace too small");
    bool any_cluster = false, any_sw = false;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r], count = schedule[2 * r + 1];
        CM_REQUIRE(code == CODE_METROPOLIS || code == CODE_HEATBATH || code == CODE_OVERRELAX ||
                       code == CODE_WOLFF || code == CODE_SWENDSEN_WANG,
                   "update code %lld has no O(n) GPU kernel", (long long)code);
        CM_REQUIRE(count >= 1 && count <= (1 << 20), "repetition count out of range");
        any_cluster |= code == CODE_WOLFF || code == CODE_SWENDSEN_WANG;
        any_sw |= code == CODE_SWENDSEN_WANG;
    }
    OnClArgs ca{};
    if (any_cluster) {
        CM_REQUIRE(n_meas_rep == nullptr, "cluster updates take no per-chain measurement caps");
        CM_REQUIRE(g.N * ndim < ((int64_t)1 << 31), "lattice too large for cluster updates");
        const OnClLayout lay = on_cl_layout(g.N, n_rep);
        CM_REQUIRE(cluster_ws != nullptr && cluster_ws_bytes >= lay.total,
                   "cluster updates need a cluster workspace of cm_on_cluster_workspace_bytes bytes");
        char* cw = (char*)cluster_ws;
        ca.spins = (float*)spins;
        ca.betas = betas_dev;
        ca.keys = keys;
        ca.sids = stream_ids;
        ca.sweeps = sweeps;
        ca.parent = (int32_t*)(cw + lay.parent);
        ca.size = (int32_t*)(cw + lay.size);
        ca.queue = (int32_t*)(cw + lay.queue);
        ca.mark = (int32_t*)(cw + lay.mark);
        ca.proj = (double*)(cw + lay.proj);
        ca.stats = (int64_t*)(cw + lay.stats);
        ca.wsize = (int64_t*)(cw + lay.wsize);
        ca.g = on_clgeo(ndim, dims);
        ca.R = n_rep;
    }
    if (n_meas == 0) return CM_OK;
    hipStream_t st = (hipStream_t)stream;
    const int nchunk = (int)((g.N + ON_CHUNK - 1) / ON_CHUNK);
    double* part = (double*)workspace;
    const int64_t total = (int64_t)n_rep * n_meas;
    hipLaunchKernelGGL(fill_i32_on, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 4096)),
                       dim3(256), 0, st, acc_out, total);
    OnArgs a{(float*)spins, betas_dev, steps_dev, keys, stream_ids, sweeps, acc_out, n_meas_rep,
             g, n_rep, n_meas, 0, 0, 0, 0};
    const int64_t rep_bytes = g.N * n * 4;
    if (!any_cluster && aux_out == nullptr &&
        (rep_bytes <= ON_BATCHED_LDS_ANY || (rep_bytes <= ON_BATCHED_LDS && n_rep <= on_cu_count()))) {
        // the whole call in one launch, each replica in LDS
        Sched sc{};
        sc.n = n_rows;
        for (int r = 0; r < n_rows; ++r) {
            sc.code[r] = (int32_t)schedule[2 * r];
            sc.count[r] = (int32_t)schedule[2 * r + 1];
        }
        const size_t lds = (size_t)g.N * n * 4;
        const dim3 bgrid((unsigned)std::min(n_rep, 65535));
        bool hb = false;
        for (int r = 0; r < n_rows; ++r) hb |= schedule[2 * r] == CODE_HEATBATH;
        if (hb) CM_HIP_CHECK(launch_batched_on<true>(n, g.d, bgrid, lds, st, a, sc, e_out, m_out, sc_out, ss_out));
        else CM_HIP_CHECK(launch_batched_on<false>(n, g.d, bgrid, lds, st, a, sc, e_out, m_out, sc_out, ss_out));
        int64_t per_meas = 0;
        for (int r = 0; r < n_rows; ++r) per_meas += schedule[2 * r + 1];
        hipLaunchKernelGGL(on_advance_sweeps, dim3((n_rep + 255) / 256), dim3(256), 0, st, sweeps,
                           n_rep, (uint64_t)per_meas, n_meas, n_meas_rep);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    }
    // When the schedule ends with a local row, its last half sweep (colour 1)
    // writes renormalised spins and the measurement renormalises colour 0.
    const int64_t last_code = schedule[2 * (n_rows - 1)];
    const bool fused = last_code == CODE_METROPOLIS || last_code == CODE_HEATBATH ||
                       last_code == CODE_OVERRELAX;
    uint64_t sw = 0;
    const int rb = (n_rep + 255) / 256;
    if (any_cluster)
        hipLaunchKernelGGL(on_cl_zero64, dim3((2 * n_rep + 255) / 256), dim3(256), 0, st, ca.stats, 2 * n_rep);
    const int or_rows = on_or_tile_rows(g, n);
    if (or_rows > 0) {
        const size_t lds = (size_t)(2 * or_rows + 4 * ON_OR_FUSE - 2) * g.M * 8;
        CM_HIP_CHECK(hipFuncSetAttribute((const void*)on_or_fused<ON_OR_FUSE>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    const OnWs wl = on_ws_layout(g.N, n_rep, n, g.d);
    // the fused over-relaxation launch writes out of place: `cur` holds the spins between launches
    float* const home = (float*)spins;
    float* const alt = n == 2 && g.d == 2 ? (float*)((char*)workspace + wl.alt) : nullptr;
    float* cur = home;
    auto other = [&](const float* p) { return p == home ? alt : home; };
    // Measurements carried by the next sweep: when a measurement ends with a
    // plain colour-1 half sweep and the next one starts with a plain colour-0
    // half sweep, that colour-0 launch (on_c0_sweep_measure) renormalises and
    // measures the previous state and the colour-1 launch after it adds the
    // chunk partials; only the call's last measurement takes the separate
    // measure launches.
    const int64_t first_code = schedule[0], first_count = schedule[1];
    const int64_t last_count = schedule[2 * (n_rows - 1) + 1];
    const bool first_plain = !(first_code == CODE_OVERRELAX && or_rows > 0 && first_count >= 2);
    const bool last_plain = !(last_code == CODE_OVERRELAX && or_rows > 0 && last_count % 2 == 0);
    const bool defer = fused && !any_cluster && first_plain && last_plain;
    const int nc0 = (int)((g.half + ON_CHUNK - 1) / ON_CHUNK);
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        a.meas = meas;
        int meas_tiles = 0;  // > 0: the fused over-relaxation launch measured
        const bool carry = defer && meas > 0;  // this measurement's first sweep carries the previous one
        if (any_cluster) hipLaunchKernelGGL(on_cl_zero64, dim3(rb), dim3(256), 0, st, ca.wsize, n_rep);
        for (int r = 0; r < n_rows; ++r) {
            const int code = (int)schedule[2 * r];
            if (code == CODE_WOLFF || code == CODE_SWENDSEN_WANG) {
                ca.sweep_add = sw;
                ca.spins = cur;
                if (n == 2) on_cluster_row<2>(ca, code, schedule[2 * r + 1], st);
                else on_cluster_row<3>(ca, code, schedule[2 * r + 1], st);
                sw += (uint64_t)schedule[2 * r + 1];
                continue;
            }
            const int64_t count = schedule[2 * r + 1];
            for (int64_t rp = 0; rp < count;) {
                a.sweep_add = sw;
                if (code == CODE_OVERRELAX && or_rows > 0 && count - rp >= 2) {
                    // two over-relaxation sweeps in one launch, tiles of rows in LDS
                    a.norm_out = fused && r == n_rows - 1 && rp + 2 == count;
                    const int tiles = (g.L1 + or_rows - 1) / or_rows;
                    const size_t lds = (size_t)(2 * or_rows + 4 * ON_OR_FUSE - 2) * g.M * 8;
                    // the measurement's last launch also measures, one partial per tile
                    // (the workspace holds nchunk >= tiles partials per replica)
                    double* tpart = a.norm_out && tiles <= nchunk ? part : nullptr;
                    if (tpart) meas_tiles = tiles;
                    a.spins = other(cur);
                    hipLaunchKernelGGL((on_or_fused<ON_OR_FUSE>), dim3((unsigned)(tiles * n_rep)),
                                       dim3(1024), lds, st, a, (const float*)cur, or_rows, tpart);
                    cur = a.spins;
                    sw += 2;
                    rp += 2;
                    continue;
                }
                for (int c = 0; c < 2; ++c) {
                    a.colour = c;
                    a.norm_out = fused && r == n_rows - 1 && rp == count - 1 && c == 1;
                    const bool first = carry && r == 0 && rp == 0;
                    if (first && c == 0) {  // measures the previous state, then updates colour 0
                        if (n == 2) launch_c0_measure<2>(code, a, part, nc0, meas - 1, st);
                        else launch_c0_measure<3>(code, a, part, nc0, meas - 1, st);
                        continue;
                    }
                    OnFinal fin{};
                    if (first)  // colour 1 adds the previous measurement's partials up
                        fin = OnFinal{part, meas - 1, nc0, e_out, m_out, sc_out, ss_out};
                    if (n == 2) launch_half_z<2>(code, a, st, fin);
                    else launch_half_z<3>(code, a, st, fin);
                }
                ++sw;
                ++rp;
            }
        }
        const int64_t sites = (int64_t)n_rep * g.N;
        const int nb = (int)std::min<int64_t>((sites + 255) / 256, 8192);
        if (meas_tiles > 0) {
            hipLaunchKernelGGL((on_measure_final<2, 2>), dim3((n_rep * 7 + 3) / 4), dim3(256), 0, st,
                               part, meas_tiles, n_rep, n_meas, meas, e_out, m_out, sc_out, ss_out, n_meas_rep);
        } else if (defer && meas + 1 < n_meas) {
            // measured by the next measurement's first sweep
        } else if (fused) {
            if (n == 2) launch_measure_c0<2>(cur, g, n_rep, part, n_meas, meas, e_out, m_out, sc_out, ss_out, n_meas_rep, st);
            else launch_measure_c0<3>(cur, g, n_rep, part, n_meas, meas, e_out, m_out, sc_out, ss_out, n_meas_rep, st);
        } else if (n == 2) {
            hipLaunchKernelGGL(on_normalize<2>, dim3(nb), dim3(256), 0, st, cur, sites, g.N, n_meas_rep, meas);
            launch_measure<2>((const float*)cur, g, n_rep, part, nchunk, n_meas, meas, e_out, m_out, sc_out, ss_out, n_meas_rep, st);
        } else {
            hipLaunchKernelGGL(on_normalize<3>, dim3(nb), dim3(256), 0, st, cur, sites, g.N, n_meas_rep, meas);
            launch_measure<3>((const float*)cur, g, n_rep, part, nchunk, n_meas, meas, e_out, m_out, sc_out, ss_out, n_meas_rep, st);
        }
        if (aux_out)
            hipLaunchKernelGGL(on_cl_record, dim3(rb), dim3(256), 0, st, n_rep, n_meas, meas,
                               acc_out, ca.wsize, ca.stats, any_sw ? 1 : 0, aux_out);
    }
    if (cur != home)  // the last fused over-relaxation launch wrote the second buffer
        hipLaunchKernelGGL(on_copy_back, dim3((unsigned)std::min<int64_t>((g.N * n / 4 + 255) / 256, 256),
                                              (unsigned)n_rep),
                           dim3(256), 0, st, home, (const float*)cur, n_rep, g.N * n, n_meas_rep);
    int64_t per_meas = 0;
    for (int r = 0; r < n_rows; ++r) per_meas += schedule[2 * r + 1];
    hipLaunchKernelGGL(on_advance_sweeps, dim3((n_rep + 255) / 256), dim3(256), 0, st, sweeps,
                       n_rep, (uint64_t)per_meas, n_meas, n_meas_rep);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // extern "C"
