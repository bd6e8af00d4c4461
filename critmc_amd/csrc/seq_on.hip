// Sequential O(n) chains (n = 2, 3) in FP64, one lane per chain, on any
// lattice given by its neighbour table nbr[N][z] in pair form (column 2a =
// +e_a, 2a+1 = -e_a; critmc/lattice.py:129-134, 173-213), in the reference's
// own site orders (critmc/kernels/on.py):
//   code 0  metropolis          sites 0..N-1 in turn (:112-137)
//   code 1  metropolis-random   N attempts at uniformly drawn sites (:112-137)
//   code 2  heatbath            sites 0..N-1 in turn (:186-261)
//   code 5  overrelax           sites 0..N-1 in turn (:140-183)
//   code 3  wolff, 4  sw        the embedded Ising model along a random axis
//                               (:264-345), neighbours from the table
// After each measurement's rows every spin is renormalised and E, M[n] and the
// stiffness sums c_a, s_a are summed in FP64 over the forward bonds in the
// reference's order: axis-major, sites ascending (:35-87, :404-413).
//
// Uniforms: attempt t of a local sweep draws Philox block(s) at (sweep, field =
// update code, TAG_SEQ, the chain's stream):
//   metropolis, n = 2: block t: w0 -> site floor(w0 N / 2^32) (random order),
//                      (w1, w2) -> Box-Muller g0, g1, w3 = acceptance
//   metropolis, n = 3: block 2t: (w0, w1) -> g0, g1, (w2, w3) -> g2; block 2t+1:
//                      w0 = acceptance, w1 -> site
//   Box-Muller:        u1 = (x+1)/2^32, u2 = x/2^32, g = sqrt(-2 ln u1) (cos, sin)(2 pi u2)
//   acceptance:        dE <= 0 or x 2^-32 < det_exp(-beta dE)
//   heat bath, n = 2:  Best & Fisher; attempt a < 32 of site t draws block t at
//                      field 2 | a << 3: w0 -> z = cos(pi w0 2^-32), w1 -> u2 =
//                      (w1+1) 2^-32, top bit of w2 -> sign; the last proposal is
//                      kept if all reject; kappa < 1e-8: phi = pi (2 w0 2^-32 - 1)
//   heat bath, n = 3:  block t: r = (w0+1) 2^-32, azimuth 2 pi w1 2^-32;
//                      |h| = 0: z = 2 w2 2^-32 - 1, azimuth 2 pi w3 2^-32
//   cluster rows:      tag TAG_CLUSTER: axis from block 0 of field 3, bond
//                      b = a N + i from word b&3 of block b>>2 (field 0), SW
//                      cluster flips from its smallest site rt (field 1), Wolff
//                      seed from word 0 of field 2
// Every update takes one value of the chain's sweep counter.  The arithmetic
// is the CPU replay's, operation for operation; only the device's libm
// (log, cos, sin, atan2, acos, exp) may differ from the host's in the last
// bits, so spins agree with the replay to rounding, not bit for bit.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cluster_core.hpp"
#include "common.hpp"
#include "det_exp.hpp"
#include "philox.hpp"
#include "seq_core.hpp"

namespace cm {

constexpr int SQO_HB_ATTEMPTS = 32;
constexpr double SQO_PI = 3.14159265358979323846;

struct SeqOnArgs {
    double* spins;           // [R][N][n]
    const int32_t* nbr;      // [N][z]
    int64_t N;
    int z, R;
    const double* betas;     // [R]
    const double* steps;     // [R]
    const uint32_t* keys;    // [R][2]
    const uint32_t* sids;    // [R]
    uint64_t* sweeps;        // [R], advanced by the kernel
    int32_t* work;           // [R][3N]: union-find parent and size, or Wolff queue and mark
    double* proj;            // [R][N]
    SeqRows rows;
    int64_t n_meas;
    double* e_out;           // [R][n_meas]
    double* m_out;           // [R][n_meas][n]
    double* c_out;           // [R][n_meas][z/2]
    double* s_out;           // [R][n_meas][z/2]
    int64_t* aux_out;        // [R][n_meas][4]
};

struct SqoChain {
    uint32_t k0, k1, sid;
    uint64_t sweep;
};

__device__ __forceinline__ u32x4 sqo_block(const SqoChain& ch, uint32_t field, uint32_t t)
{
    return philox4x32_10(u32x4{t, (uint32_t)ch.sweep, ctr_z(ch.sweep, field, TAG_SEQ), ch.sid}, ch.k0, ch.k1);
}

__device__ __forceinline__ void sqo_bm(uint32_t xa, uint32_t xb, double& g0, double& g1)
{
    const double u1 = ((double)xa + 1.0) * 0x1p-32, u2 = (double)xb * 0x1p-32;
    const double r = sqrt(-2.0 * log(u1));
    g0 = r * cos(2.0 * SQO_PI * u2);
    g1 = r * sin(2.0 * SQO_PI * u2);
}

template <int NC>
__device__ __forceinline__ void sqo_field(const double* __restrict__ s, const int32_t* __restrict__ nbr,
                                          int z, int64_t i, double (&h)[NC])
{
#pragma unroll
    for (int c = 0; c < NC; ++c) h[c] = 0.0;
    for (int j = 0; j < z; ++j) {
        const double* sj = s + (int64_t)nbr[i * z + j] * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) h[c] += sj[c];
    }
}

template <int NC>
__device__ int64_t sqo_metropolis(double* __restrict__ s, const SeqOnArgs& a, int code, double beta,
                                  double step, const SqoChain& ch)
{
    const int64_t N = a.N;
    int64_t acc = 0;
    for (int64_t t = 0; t < N; ++t) {
        double g[4];
        uint32_t xs, xu;
        if constexpr (NC == 2) {
            const u32x4 w = sqo_block(ch, (uint32_t)code, (uint32_t)t);
            sqo_bm(w.y, w.z, g[0], g[1]);
            xs = w.x;
            xu = w.w;
        } else {
            const u32x4 w = sqo_block(ch, (uint32_t)code, (uint32_t)(2 * t));
            const u32x4 v = sqo_block(ch, (uint32_t)code, (uint32_t)(2 * t + 1));
            sqo_bm(w.x, w.y, g[0], g[1]);
            sqo_bm(w.z, w.w, g[2], g[3]);
            xu = v.x;
            xs = v.y;
        }
        const int64_t i = code == 1 ? (int64_t)(((uint64_t)xs * (uint64_t)N) >> 32) : t;
        double h[NC], tr[NC], nrm = 0.0, de = 0.0;
        sqo_field<NC>(s, a.nbr, a.z, i, h);
        double* si = s + i * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            tr[c] = si[c] + step * g[c];
            nrm += tr[c] * tr[c];
        }
        nrm = sqrt(nrm);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            tr[c] /= nrm;
            de -= (tr[c] - si[c]) * h[c];
        }
        if (de <= 0.0 || (double)xu * 0x1p-32 < det_exp(-beta * de)) {
#pragma unroll
            for (int c = 0; c < NC; ++c) si[c] = tr[c];
            ++acc;
        }
    }
    return acc;
}

// Best & Fisher von Mises deviate (critmc/kernels/on.py:186-204).
__device__ double sqo_von_mises(double kappa, const SqoChain& ch, uint32_t t)
{
    if (kappa < 1e-8) {
        const u32x4 b = sqo_block(ch, 2, t);
        return SQO_PI * (2.0 * ((double)b.x * 0x1p-32) - 1.0);
    }
    const double root = sqrt(1.0 + 4.0 * kappa * kappa);
    const double tau = 1.0 + root;
    const double rho = 2.0 * kappa * tau / ((root + 1.0) * (tau + sqrt(2.0 * tau)));
    const double r = (1.0 + rho * rho) / (2.0 * rho);
    double f = 0.0;
    uint32_t sign = 0;
    for (int att = 0; att < SQO_HB_ATTEMPTS; ++att) {
        const u32x4 b = sqo_block(ch, 2u | ((uint32_t)att << 3), t);
        const double zz = cos(SQO_PI * ((double)b.x * 0x1p-32));
        f = (1.0 + r * zz) / (r + zz);
        const double c = kappa * (r - f);
        const double u2 = ((double)b.y + 1.0) * 0x1p-32;
        sign = b.z >> 31;
        if (c * (2.0 - c) - u2 > 0.0 || log(c / u2) + 1.0 - c >= 0.0) break;
    }
    const double theta = acos(fmin(1.0, fmax(-1.0, f)));
    return sign == 0 ? theta : -theta;
}

template <int NC>
__device__ void sqo_heatbath(double* __restrict__ s, const SeqOnArgs& a, double beta, const SqoChain& ch)
{
    for (int64_t t = 0; t < a.N; ++t) {
        double h[NC];
        sqo_field<NC>(s, a.nbr, a.z, t, h);
        double* si = s + t * NC;
        double hn2 = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) hn2 += h[c] * h[c];
        const double hn = sqrt(hn2);
        const double kappa = beta * hn;
        if constexpr (NC == 2) {
            const double phi0 = hn > 0.0 ? atan2(h[1], h[0]) : 0.0;
            const double phi = phi0 + sqo_von_mises(kappa, ch, (uint32_t)t);
            si[0] = cos(phi);
            si[1] = sin(phi);
        } else {
            const u32x4 b = sqo_block(ch, 2, (uint32_t)t);
            if (hn == 0.0) {
                const double zz = 2.0 * ((double)b.z * 0x1p-32) - 1.0;
                const double az = 2.0 * SQO_PI * ((double)b.w * 0x1p-32);
                const double w = sqrt(fmax(0.0, 1.0 - zz * zz));
                si[0] = w * cos(az);
                si[1] = w * sin(az);
                si[2] = zz;
                continue;
            }
            const double r = ((double)b.x + 1.0) * 0x1p-32;
            double u = kappa < 1e-8 ? 2.0 * r - 1.0 : 1.0 + log(r + (1.0 - r) * exp(-2.0 * kappa)) / kappa;
            u = fmin(1.0, fmax(-1.0, u));
            const double hx = h[0] / hn, hy = h[1] / hn, hz = h[2] / hn;
            const double ax = fabs(hx) < 0.9 ? 1.0 : 0.0, ay = 1.0 - ax;
            const double proj = ax * hx + ay * hy;
            double e1[3] = {ax - proj * hx, ay - proj * hy, -proj * hz};
            const double nrm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
            for (int c = 0; c < 3; ++c) e1[c] /= nrm;
            const double e2[3] = {hy * e1[2] - hz * e1[1], hz * e1[0] - hx * e1[2], hx * e1[1] - hy * e1[0]};
            const double phi = 2.0 * SQO_PI * ((double)b.y * 0x1p-32);
            const double w = sqrt(fmax(0.0, 1.0 - u * u));
            const double cp = cos(phi), sp = sin(phi);
            si[0] = u * hx + w * (cp * e1[0] + sp * e2[0]);
            si[1] = u * hy + w * (cp * e1[1] + sp * e2[1]);
            si[2] = u * hz + w * (cp * e1[2] + sp * e2[2]);
        }
    }
}

template <int NC>
__device__ void sqo_overrelax(double* __restrict__ s, const SeqOnArgs& a)
{
    for (int64_t t = 0; t < a.N; ++t) {
        double h[NC];
        sqo_field<NC>(s, a.nbr, a.z, t, h);
        double* si = s + t * NC;
        double hh = 0.0, sh = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            hh += h[c] * h[c];
            sh += si[c] * h[c];
        }
        if (hh == 0.0) continue;
        const double f = 2.0 * sh / hh;
#pragma unroll
        for (int c = 0; c < NC; ++c) si[c] = f * h[c] - si[c];
    }
}

__device__ __forceinline__ uint32_t sqo_cl_word(const SqoChain& ch, uint32_t field, int64_t index)
{
    const u32x4 w = philox4x32_10(
        u32x4{(uint32_t)(index >> 2), (uint32_t)ch.sweep, ctr_z(ch.sweep, field, TAG_CLUSTER), ch.sid},
        ch.k0, ch.k1);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    return ws[index & 3];
}

template <int NC>
__device__ void sqo_axis(const SqoChain& ch, double (&r)[NC])
{
    const u32x4 w = philox4x32_10(u32x4{0u, (uint32_t)ch.sweep, ctr_z(ch.sweep, 3, TAG_CLUSTER), ch.sid},
                                  ch.k0, ch.k1);
    if constexpr (NC == 2) {
        const double phi = 2.0 * SQO_PI * ((double)w.x * 0x1p-32);
        r[0] = cos(phi);
        r[1] = sin(phi);
    } else {
        const double zz = 2.0 * ((double)w.x * 0x1p-32) - 1.0, az = 2.0 * SQO_PI * ((double)w.y * 0x1p-32);
        const double ww = sqrt(fmax(0.0, 1.0 - zz * zz));
        r[0] = ww * cos(az);
        r[1] = ww * sin(az);
        r[2] = zz;
    }
}

template <int NC>
__device__ __forceinline__ double sqo_proj(const double* __restrict__ s, int64_t i, const double (&r)[NC])
{
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) p += s[i * NC + c] * r[c];
    return p;
}

__device__ __forceinline__ bool sqo_active(double x, double beta, uint32_t w)
{
    return x > 0.0 && (double)w * 0x1p-32 < 1.0 - det_exp(-2.0 * beta * x);
}

__device__ __forceinline__ int32_t sqo_find(int32_t* parent, int32_t i)
{
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
}

template <int NC>
__device__ void sqo_sw(double* __restrict__ s, const SeqOnArgs& a, double beta, const SqoChain& ch,
                       int32_t* parent, int32_t* size, double* proj, int64_t& c2, int64_t& ncl)
{
    const int64_t N = a.N;
    double r[NC];
    sqo_axis<NC>(ch, r);
    for (int64_t i = 0; i < N; ++i) {
        parent[i] = (int32_t)i;
        proj[i] = sqo_proj<NC>(s, i, r);
    }
    for (int ax = 0; ax < a.z / 2; ++ax)
        for (int64_t i = 0; i < N; ++i) {
            const int64_t j = a.nbr[i * a.z + 2 * ax];
            if (!sqo_active(proj[i] * proj[j], beta, sqo_cl_word(ch, 0, (int64_t)ax * N + i))) continue;
            const int32_t ri = sqo_find(parent, (int32_t)i), rj = sqo_find(parent, (int32_t)j);
            if (ri == rj) continue;
            if (ri < rj) parent[rj] = ri;
            else parent[ri] = rj;
        }
    for (int64_t i = 0; i < N; ++i) size[i] = 0;
    for (int64_t i = 0; i < N; ++i) size[sqo_find(parent, (int32_t)i)] += 1;
    c2 = 0;
    ncl = 0;
    for (int64_t i = 0; i < N; ++i) {
        const int32_t rt = sqo_find(parent, (int32_t)i);
        if (rt == i) {
            c2 += (int64_t)size[i] * size[i];
            ncl += 1;
        }
        if ((sqo_cl_word(ch, 1, rt) >> 31) == 0)
#pragma unroll
            for (int c = 0; c < NC; ++c) s[i * NC + c] -= 2.0 * proj[i] * r[c];
    }
}

template <int NC>
__device__ int64_t sqo_wolff(double* __restrict__ s, const SeqOnArgs& a, double beta, const SqoChain& ch,
                             int32_t* queue, int32_t* mark, int tag, double* proj)
{
    const int64_t N = a.N;
    double r[NC];
    sqo_axis<NC>(ch, r);
    const int64_t seed = (int64_t)(((uint64_t)sqo_cl_word(ch, 2, 0) * (uint64_t)N) >> 32);
    const double p0 = sqo_proj<NC>(s, seed, r);
    proj[seed] = p0;
#pragma unroll
    for (int c = 0; c < NC; ++c) s[seed * NC + c] -= 2.0 * p0 * r[c];
    mark[seed] = tag;
    queue[0] = (int32_t)seed;
    int64_t head = 0, tail = 1;
    while (head < tail) {
        const int64_t k = queue[head++];
        const double pk = proj[k];
        for (int cc = 0; cc < a.z; ++cc) {
            const int64_t j = a.nbr[k * a.z + cc];
            if (mark[j] == tag) continue;
            const double pj = sqo_proj<NC>(s, j, r);
            const int64_t b = (int64_t)(cc >> 1) * N + ((cc & 1) ? j : k);
            if (!sqo_active(pk * pj, beta, sqo_cl_word(ch, 0, b))) continue;
            proj[j] = pj;
#pragma unroll
            for (int c = 0; c < NC; ++c) s[j * NC + c] -= 2.0 * pj * r[c];
            mark[j] = tag;
            queue[tail++] = (int32_t)j;
        }
    }
    return tail;
}

template <int NC>
__device__ void sqo_measure(double* __restrict__ s, const SeqOnArgs& a, double* e, double* m, double* sc,
                            double* ss)
{
    const int64_t N = a.N;
    const int z = a.z;
    for (int64_t i = 0; i < N; ++i) {
        double n2 = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) n2 += s[i * NC + c] * s[i * NC + c];
        const double nrm = sqrt(n2);
#pragma unroll
        for (int c = 0; c < NC; ++c) s[i * NC + c] /= nrm;
    }
    double ee = 0.0;
    for (int ax = 0; ax < z / 2; ++ax)
        for (int64_t i = 0; i < N; ++i) {
            const int64_t j = a.nbr[i * z + 2 * ax];
            double dot = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) dot += s[i * NC + c] * s[j * NC + c];
            ee -= dot;
        }
    *e = ee;
    double mm[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) mm[c] = 0.0;
    for (int64_t i = 0; i < N; ++i)
#pragma unroll
        for (int c = 0; c < NC; ++c) mm[c] += s[i * NC + c];
#pragma unroll
    for (int c = 0; c < NC; ++c) m[c] = mm[c];
    for (int ax = 0; ax < z / 2; ++ax) {
        double cs = 0.0, sn = 0.0;
        for (int64_t i = 0; i < N; ++i) {
            const int64_t j = a.nbr[i * z + 2 * ax];
            cs += s[i * NC] * s[j * NC] + s[i * NC + 1] * s[j * NC + 1];
            sn += s[i * NC] * s[j * NC + 1] - s[i * NC + 1] * s[j * NC];
        }
        sc[ax] = cs;
        ss[ax] = sn;
    }
}

template <int NC>
__global__ __launch_bounds__(64) void seq_on_run(SeqOnArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    const int64_t N = a.N;
    double* s = a.spins + (int64_t)r * N * NC;
    int32_t* work = a.work + (int64_t)r * 3 * N;
    double* proj = a.proj + (int64_t)r * N;
    SqoChain ch{a.keys[2 * r], a.keys[2 * r + 1], a.sids[r], a.sweeps[r]};
    const double beta = a.betas[r], step = a.steps[r];
    const int A = a.z / 2;
    for (int64_t meas = 0; meas < a.n_meas; ++meas) {
        int64_t acc = 0, wsize = 0, c2 = 0, ncl = 0;
        for (int row = 0; row < a.rows.n; ++row) {
            const int code = a.rows.code[row];
            if (code == CODE_WOLFF)
                for (int64_t i = 0; i < N; ++i) work[2 * N + i] = -1;
            for (int rp = 0; rp < a.rows.count[row]; ++rp) {
                if (code == CODE_METROPOLIS || code == CODE_METROPOLIS_RANDOM)
                    acc += sqo_metropolis<NC>(s, a, code, beta, step, ch);
                else if (code == CODE_HEATBATH)
                    sqo_heatbath<NC>(s, a, beta, ch);
                else if (code == CODE_OVERRELAX)
                    sqo_overrelax<NC>(s, a);
                else if (code == CODE_WOLFF)
                    wsize += sqo_wolff<NC>(s, a, beta, ch, work, work + 2 * N, rp, proj);
                else
                    sqo_sw<NC>(s, a, beta, ch, work, work + N, proj, c2, ncl);
                ++ch.sweep;
            }
        }
        const int64_t o = (int64_t)r * a.n_meas + meas;
        sqo_measure<NC>(s, a, a.e_out + o, a.m_out + o * NC, a.c_out + o * A, a.s_out + o * A);
        a.aux_out[4 * o + 0] = acc;
        a.aux_out[4 * o + 1] = wsize;
        a.aux_out[4 * o + 2] = c2;
        a.aux_out[4 * o + 3] = ncl;
    }
    a.sweeps[r] = ch.sweep;
}

static int64_t sqo_ws_layout(int64_t N, int32_t R, int64_t* off_work, int64_t* off_proj)
{
    const int64_t tabs = ((int64_t)R * 16 + 255) / 256 * 256;
    const int64_t work = ((int64_t)R * 3 * N * 4 + 255) / 256 * 256;
    if (off_work) *off_work = tabs;
    if (off_proj) *off_proj = tabs + work;
    return tabs + work + ((int64_t)R * N * 8 + 255) / 256 * 256;
}

}  // namespace cm

using namespace cm;

extern "C" {

int64_t cm_on_table_workspace_bytes(int64_t n_sites, int32_t n_rep)
{
    return sqo_ws_layout(n_sites, n_rep, nullptr, nullptr);
}

int cm_on_table_run(void* spins, int64_t n_sites, int32_t z, const int32_t* nbr, int32_t n_rep, int32_t n,
                    const double* betas, const double* steps, const uint32_t* keys,
                    const uint32_t* stream_ids, uint64_t* sweeps, const int64_t* schedule, int32_t n_rows,
                    int64_t n_meas, double* e_out, double* m_out, double* c_out, double* s_out,
                    int64_t* aux_out, void* workspace, int64_t workspace_bytes, void* stream)
{
    CM_REQUIRE(n == 2 || n == 3, "the O(n) kernels support n = 2 and 3, got %d", n);
    CM_REQUIRE(n_sites >= 1 && n_sites < ((int64_t)1 << 28), "lattice size out of range");
    CM_REQUIRE(z >= 2 && z <= SEQ_ZMAX && z % 2 == 0, "need 2..%d neighbour columns in pairs", SEQ_ZMAX);
    CM_REQUIRE(n_rep >= 1, "n_rep must be positive");
    CM_REQUIRE(n_meas >= 0, "n_meas must not be negative");
    CM_REQUIRE(n_rows >= 1 && n_rows <= 8, "schedule must have 1..8 rows");
    CM_REQUIRE(workspace_bytes >= cm_on_table_workspace_bytes(n_sites, n_rep), "workspace too small");
    SeqRows rows{};
    rows.n = n_rows;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r], count = schedule[2 * r + 1];
        CM_REQUIRE(code == CODE_METROPOLIS || code == CODE_METROPOLIS_RANDOM || code == CODE_HEATBATH ||
                       code == CODE_OVERRELAX || code == CODE_WOLFF || code == CODE_SWENDSEN_WANG,
                   "update code %lld has no O(n) kernel", (long long)code);
        CM_REQUIRE(count >= 1 && count <= (1 << 20), "repetition count out of range");
        rows.code[r] = (int32_t)code;
        rows.count[r] = (int32_t)count;
    }
    for (int r = 0; r < n_rep; ++r)
        CM_REQUIRE(std::isfinite(betas[r]) && std::isfinite(steps[r]), "beta and step must be finite");
    if (n_meas == 0) return CM_OK;
    hipStream_t st = (hipStream_t)stream;
    int64_t o_work, o_proj;
    sqo_ws_layout(n_sites, n_rep, &o_work, &o_proj);
    double* dtab = (double*)workspace;
    std::vector<double> tab(2 * (size_t)n_rep);
    std::copy(betas, betas + n_rep, tab.begin());
    std::copy(steps, steps + n_rep, tab.begin() + n_rep);
    CM_HIP_CHECK(upload_async(dtab, tab.data(), tab.size() * 8, st));
    SeqOnArgs a{};
    a.spins = (double*)spins;
    a.nbr = nbr;
    a.N = n_sites;
    a.z = z;
    a.R = n_rep;
    a.betas = dtab;
    a.steps = dtab + n_rep;
    a.keys = keys;
    a.sids = stream_ids;
    a.sweeps = sweeps;
    a.work = (int32_t*)((char*)workspace + o_work);
    a.proj = (double*)((char*)workspace + o_proj);
    a.rows = rows;
    a.n_meas = n_meas;
    a.e_out = e_out;
    a.m_out = m_out;
    a.c_out = c_out;
    a.s_out = s_out;
    a.aux_out = aux_out;
    const dim3 grid((unsigned)((n_rep + 63) / 64));
    if (n == 2) hipLaunchKernelGGL(seq_on_run<2>, grid, dim3(64), 0, st, a);
    else hipLaunchKernelGGL(seq_on_run<3>, grid, dim3(64), 0, st, a);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // extern "C"
