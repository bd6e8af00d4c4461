/*
 * ORACLE (test infrastructure only) -- FP64 replay of the checkerboard O(n)
 * schedule (n = 2, 3) with the GPU path's uniforms.  The GPU runs in FP32, so
 * agreement is to a stated tolerance, not bit-exact.
 *
 * Restated from the reference (critmc/kernels/on.py):
 *   Metropolis :112-137, over-relaxation :140-183, normalize :78-87,
 *   energy / magnetization / stiffness_sums :35-76, run loop :348-418.
 * Heat bath (:186-261) in FP64, uniforms at tag ORC_TAG_ONHB: XY attempt a of
 * colour site k draws block k with colour field colour | a << 1 (word 0 ->
 * z = cos(pi U), word 1 -> u2 = (x+1)/2^32, top bit of word 2 -> sign), at most
 * 128 attempts, the last proposal kept if all reject; kappa < 1e-8 gives
 * phi = pi (2U - 1) from word 0 of attempt 0.  Heisenberg: block k, r = (x0+1)/2^32,
 * azimuth 2 pi x1/2^32; |h| = 0: z = 2 x2/2^32 - 1, azimuth from x3.
 * Cluster updates (:264-345): the embedded Ising model along a random unit
 * vector r (block 0 of colour field 3, tag ORC_TAG_CLUSTER: XY angle 2 pi U from
 * word 0; n = 3: z = 2U - 1 from word 0, azimuth from word 1); p_i = S_i . r;
 * bond b = a N + i (i, +e_a neighbour j) active iff x = p_i p_j > 0 and
 * u < 1 - det_exp(-2 beta x), u = word b&3 of block b>>2 (colour field 0);
 * Swendsen-Wang reflects S -> S - 2 p r the clusters whose smallest site rt
 * has word rt&3 of block rt>>2 (colour field 1) with top bit clear; Wolff
 * grows from seed floor(x N / 2^32) (word 0 of block 0, colour field 2) with
 * p_k the member's projection before its reflection.  One sweep-counter value
 * per update.
 * Uniforms: colour site k uses Philox block k (n = 2) or blocks 2k, 2k+1
 * (n = 3) at (sweep, colour, TAG_ON, stream); Box-Muller with
 * u1 = (x+1)/2^32, u2 = x/2^32: g0 = sqrt(-2 ln u1) cos(2 pi u2),
 * g1 = ... sin(2 pi u2); the acceptance uniform is word 2 of block k (n = 2) or
 * word 0 of block 2k+1 (n = 3).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "cb.h"
#include "philox.h"

static void orc_bm(uint32_t xa, uint32_t xb, double *g0, double *g1)
{
    double u1 = ((double)xa + 1.0) * 0x1p-32, u2 = (double)xb * 0x1p-32;
    double r = sqrt(-2.0 * log(u1));
    *g0 = r * cos(2.0 * M_PI * u2);
    *g1 = r * sin(2.0 * M_PI * u2);
}

static double orc_u0(uint32_t x) { return ((double)x + 1.0) * 0x1p-32; }
static double orc_u1(uint32_t x) { return (double)x * 0x1p-32; }

/* Best & Fisher von Mises deviate (critmc/kernels/on.py:186-204). */
static double orc_von_mises(double kappa, uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep,
                            int colour, uint32_t blk)
{
    uint32_t b[4];
    if (kappa < 1e-8) {
        orc_draw_block(k0, k1, sid, sweep, colour, ORC_TAG_ONHB, blk, b);
        return M_PI * (2.0 * orc_u1(b[0]) - 1.0);
    }
    double root = sqrt(1.0 + 4.0 * kappa * kappa);
    double tau = 1.0 + root;
    double rho = 2.0 * kappa * tau / ((root + 1.0) * (tau + sqrt(2.0 * tau)));
    double r = (1.0 + rho * rho) / (2.0 * rho);
    double f = 0.0;
    uint32_t sign = 0;
    for (int a = 0; a < 128; ++a) {
        orc_draw_block(k0, k1, sid, sweep, colour | (a << 1), ORC_TAG_ONHB, blk, b);
        double z = cos(M_PI * orc_u1(b[0]));
        f = (1.0 + r * z) / (r + z);
        double c = kappa * (r - f);
        double u2 = orc_u0(b[1]);
        sign = b[2] >> 31;
        if (c * (2.0 - c) - u2 > 0.0 || log(c / u2) + 1.0 - c >= 0.0) break;
    }
    double theta = acos(fmin(1.0, fmax(-1.0, f)));
    return sign == 0 ? theta : -theta;
}

/* Conditional sample of spin si given field h (critmc/kernels/on.py:207-261). */
static void orc_on_heatbath_site(double *si, const double *h, int n, double beta, uint32_t k0,
                                 uint32_t k1, uint32_t sid, uint64_t sweep, int colour, uint32_t k)
{
    double hn2 = 0.0;
    for (int c = 0; c < n; ++c) hn2 += h[c] * h[c];
    double hn = sqrt(hn2);
    double kappa = beta * hn;
    if (n == 2) {
        double phi0 = hn > 0.0 ? atan2(h[1], h[0]) : 0.0;
        double phi = phi0 + orc_von_mises(kappa, k0, k1, sid, sweep, colour, k);
        si[0] = cos(phi);
        si[1] = sin(phi);
        return;
    }
    uint32_t b[4];
    orc_draw_block(k0, k1, sid, sweep, colour, ORC_TAG_ONHB, k, b);
    if (hn == 0.0) {
        double z = 2.0 * orc_u1(b[2]) - 1.0, az = 2.0 * M_PI * orc_u1(b[3]);
        double w = sqrt(fmax(0.0, 1.0 - z * z));
        si[0] = w * cos(az);
        si[1] = w * sin(az);
        si[2] = z;
        return;
    }
    double r = orc_u0(b[0]);
    double u = kappa < 1e-8 ? 2.0 * r - 1.0 : 1.0 + log(r + (1.0 - r) * exp(-2.0 * kappa)) / kappa;
    u = fmin(1.0, fmax(-1.0, u));
    double hx = h[0] / hn, hy = h[1] / hn, hz = h[2] / hn;
    double ax = fabs(hx) < 0.9 ? 1.0 : 0.0, ay = 1.0 - ax;
    double proj = ax * hx + ay * hy;
    double e1[3] = {ax - proj * hx, ay - proj * hy, -proj * hz};
    double nrm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int c = 0; c < 3; ++c) e1[c] /= nrm;
    double e2[3] = {hy * e1[2] - hz * e1[1], hz * e1[0] - hx * e1[2], hx * e1[1] - hy * e1[0]};
    double phi = 2.0 * M_PI * orc_u1(b[1]);
    double w = sqrt(fmax(0.0, 1.0 - u * u));
    double cp = cos(phi), sp = sin(phi);
    si[0] = u * hx + w * (cp * e1[0] + sp * e2[0]);
    si[1] = u * hy + w * (cp * e1[1] + sp * e2[1]);
    si[2] = u * hz + w * (cp * e1[2] + sp * e2[2]);
}

/* One half sweep; returns accepted moves (Metropolis) or 0. */
int64_t orc_on_half_sweep(double *s, int ndim, const int32_t *dims, int n, int code, int colour,
                          double beta, double step, uint32_t k0, uint32_t k1, uint32_t sid,
                          uint64_t sweep)
{
    orc_cb g;
    if (orc_cb_init(&g, ndim, dims)) return -1;
    const int z = 2 * ndim;
    int64_t nb[6], acc = 0;
    for (int64_t k = 0; k < g.half; ++k) {
        int64_t i = orc_cb_site(&g, k, colour, nb);
        double h[3] = {0, 0, 0};
        for (int j = 0; j < z; ++j)
            for (int c = 0; c < n; ++c) h[c] += s[nb[j] * n + c];
        double *si = s + i * n;
        if (code == 2) {
            orc_on_heatbath_site(si, h, n, beta, k0, k1, sid, sweep, colour, (uint32_t)k);
        } else if (code == 5) {
            double hh = 0, sh = 0;
            for (int c = 0; c < n; ++c) {
                hh += h[c] * h[c];
                sh += si[c] * h[c];
            }
            if (hh == 0.0) continue;
            double f = 2.0 * sh / hh;
            for (int c = 0; c < n; ++c) si[c] = f * h[c] - si[c];
        } else {
            uint32_t b1[4], b2[4];
            double gg[4];
            uint32_t xacc;
            if (n == 2) {
                orc_draw_block(k0, k1, sid, sweep, colour, ORC_TAG_ON, (uint32_t)k, b1);
                orc_bm(b1[0], b1[1], &gg[0], &gg[1]);
                xacc = b1[2];
            } else {
                orc_draw_block(k0, k1, sid, sweep, colour, ORC_TAG_ON, (uint32_t)(2 * k), b1);
                orc_draw_block(k0, k1, sid, sweep, colour, ORC_TAG_ON, (uint32_t)(2 * k + 1), b2);
                orc_bm(b1[0], b1[1], &gg[0], &gg[1]);
                orc_bm(b1[2], b1[3], &gg[2], &gg[3]);
                xacc = b2[0];
            }
            double tr[3], nrm = 0.0, de = 0.0;
            for (int c = 0; c < n; ++c) {
                tr[c] = si[c] + step * gg[c];
                nrm += tr[c] * tr[c];
            }
            nrm = sqrt(nrm);
            for (int c = 0; c < n; ++c) {
                tr[c] /= nrm;
                de -= (tr[c] - si[c]) * h[c];
            }
            if (de <= 0.0 || (double)xacc * 0x1p-32 < exp(-beta * de)) {
                for (int c = 0; c < n; ++c) si[c] = tr[c];
                ++acc;
            }
        }
    }
    return acc;
}

double orc_det_exp(double x);

static int64_t orc_on_nbr(const int32_t *dims, int ndim, int64_t i, int a, int fwd)
{
    (void)ndim;
    int64_t stride = 1;
    for (int b = 0; b < a; ++b) stride *= dims[b];
    const int64_t L = dims[a], x = (i / stride) % L;
    if (fwd) return x + 1 == L ? i - (L - 1) * stride : i + stride;
    return x == 0 ? i + (L - 1) * stride : i - stride;
}

static uint32_t orc_on_word(uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, int field,
                            int64_t index)
{
    uint32_t w[4];
    orc_draw_block(k0, k1, sid, sweep, field, ORC_TAG_CLUSTER, (uint32_t)(index >> 2), w);
    return w[index & 3];
}

static void orc_on_axis(int n, uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, double *r)
{
    uint32_t w[4];
    orc_draw_block(k0, k1, sid, sweep, 3, ORC_TAG_CLUSTER, 0, w);
    if (n == 2) {
        double phi = 2.0 * M_PI * ((double)w[0] * 0x1p-32);
        r[0] = cos(phi);
        r[1] = sin(phi);
    } else {
        double z = 2.0 * ((double)w[0] * 0x1p-32) - 1.0, az = 2.0 * M_PI * ((double)w[1] * 0x1p-32);
        double ww = sqrt(fmax(0.0, 1.0 - z * z));
        r[0] = ww * cos(az);
        r[1] = ww * sin(az);
        r[2] = z;
    }
}

static double orc_on_proj(const double *s, int n, int64_t i, const double *r)
{
    double p = 0.0;
    for (int c = 0; c < n; ++c) p += s[i * n + c] * r[c];
    return p;
}

static int orc_on_active(double x, double beta, uint32_t w)
{
    return x > 0.0 && (double)w * 0x1p-32 < 1.0 - orc_det_exp(-2.0 * beta * x);
}

static int32_t orc_on_find(int32_t *parent, int32_t i)
{
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
}

static void orc_on_sw(double *s, int ndim, const int32_t *dims, int n, double beta, uint32_t k0,
                      uint32_t k1, uint32_t sid, uint64_t sweep, int64_t N, int32_t *parent,
                      int32_t *size, double *proj, int64_t *c2, int64_t *ncl)
{
    double r[3];
    orc_on_axis(n, k0, k1, sid, sweep, r);
    for (int64_t i = 0; i < N; ++i) {
        parent[i] = (int32_t)i;
        proj[i] = orc_on_proj(s, n, i, r);
    }
    for (int a = 0; a < ndim; ++a)
        for (int64_t i = 0; i < N; ++i) {
            int64_t j = orc_on_nbr(dims, ndim, i, a, 1);
            if (!orc_on_active(proj[i] * proj[j], beta, orc_on_word(k0, k1, sid, sweep, 0, (int64_t)a * N + i)))
                continue;
            int32_t ri = orc_on_find(parent, (int32_t)i), rj = orc_on_find(parent, (int32_t)j);
            if (ri == rj) continue;
            if (ri < rj) parent[rj] = ri;
            else parent[ri] = rj;
        }
    for (int64_t i = 0; i < N; ++i) size[i] = 0;
    for (int64_t i = 0; i < N; ++i) size[orc_on_find(parent, (int32_t)i)] += 1;
    *c2 = 0;
    *ncl = 0;
    for (int64_t i = 0; i < N; ++i) {
        int32_t rt = orc_on_find(parent, (int32_t)i);
        if (rt == i) {
            *c2 += (int64_t)size[i] * size[i];
            *ncl += 1;
        }
        if ((orc_on_word(k0, k1, sid, sweep, 1, rt) >> 31) == 0)
            for (int c = 0; c < n; ++c) s[i * n + c] -= 2.0 * proj[i] * r[c];
    }
}

static int64_t orc_on_wolff(double *s, int ndim, const int32_t *dims, int n, double beta,
                            uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, int64_t N,
                            int32_t *queue, int32_t *mark, int tag, double *proj)
{
    double r[3];
    orc_on_axis(n, k0, k1, sid, sweep, r);
    int64_t seed = (int64_t)(((uint64_t)orc_on_word(k0, k1, sid, sweep, 2, 0) * (uint64_t)N) >> 32);
    double p0 = orc_on_proj(s, n, seed, r);
    proj[seed] = p0;
    for (int c = 0; c < n; ++c) s[seed * n + c] -= 2.0 * p0 * r[c];
    mark[seed] = tag;
    queue[0] = (int32_t)seed;
    int64_t head = 0, tail = 1;
    while (head < tail) {
        int64_t k = queue[head++];
        double pk = proj[k];
        for (int cc = 0; cc < 2 * ndim; ++cc) {
            int ax = cc >> 1;
            int64_t j = orc_on_nbr(dims, ndim, k, ax, (cc & 1) == 0);
            if (mark[j] == tag) continue;
            double pj = orc_on_proj(s, n, j, r);
            int64_t b = (int64_t)ax * N + ((cc & 1) ? j : k);
            if (!orc_on_active(pk * pj, beta, orc_on_word(k0, k1, sid, sweep, 0, b))) continue;
            proj[j] = pj;
            for (int c = 0; c < n; ++c) s[j * n + c] -= 2.0 * pj * r[c];
            mark[j] = tag;
            queue[tail++] = (int32_t)j;
        }
    }
    return tail;
}

/* Normalise, then E, M[n], c_a, s_a (on.py:35-87) over forward bonds. */
int orc_on_measure(double *s, int ndim, const int32_t *dims, int n, double *e, double *m,
                   double *sc, double *ss)
{
    orc_cb g;
    if (orc_cb_init(&g, ndim, dims)) return -1;
    for (int64_t i = 0; i < g.N; ++i) {
        double n2 = 0;
        for (int c = 0; c < n; ++c) n2 += s[i * n + c] * s[i * n + c];
        double nrm = sqrt(n2);
        for (int c = 0; c < n; ++c) s[i * n + c] /= nrm;
    }
    *e = 0;
    for (int c = 0; c < n; ++c) m[c] = 0;
    for (int a = 0; a < ndim; ++a) sc[a] = ss[a] = 0;
    int64_t stride[3] = {1, g.L[0], (int64_t)g.L[0] * g.L[1]};
    for (int64_t i = 0; i < g.N; ++i) {
        for (int c = 0; c < n; ++c) m[c] += s[i * n + c];
        int64_t rest = i;
        for (int a = 0; a < ndim; ++a) {
            int x = (int)(rest % g.L[a]);
            rest /= g.L[a];
            int64_t j = i + (x + 1 == g.L[a] ? -(int64_t)(g.L[a] - 1) * stride[a] : stride[a]);
            double dot = 0;
            for (int c = 0; c < n; ++c) dot += s[i * n + c] * s[j * n + c];
            *e -= dot;
            sc[a] += s[i * n] * s[j * n] + s[i * n + 1] * s[j * n + 1];
            ss[a] += s[i * n] * s[j * n + 1] - s[i * n + 1] * s[j * n];
        }
    }
    return 0;
}

int orc_on_run_aux(double *s, int ndim, const int32_t *dims, int n, double beta, double step,
                   uint32_t k0, uint32_t k1, uint32_t sid, uint64_t *sweep, const int64_t *sched,
                   int nrows, int64_t n_meas, double *e_out, double *m_out, double *sc_out,
                   double *ss_out, int64_t *acc_out, int64_t *aux_out)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    int32_t *work = NULL;
    double *proj = NULL;
    for (int r = 0; r < nrows; ++r) {
        int code = (int)sched[2 * r];
        if (code != 0 && code != 2 && code != 3 && code != 4 && code != 5) return -3;
        if ((code == 3 || code == 4) && !work) {
            work = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)N);
            proj = (double *)malloc(sizeof(double) * (size_t)N);
        }
    }
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t acc = 0, wsize = 0, c2 = 0, ncl = 0;
        for (int r = 0; r < nrows; ++r) {
            int code = (int)sched[2 * r];
            if (code == 3) {
                for (int64_t i = 0; i < N; ++i) work[2 * N + i] = -1;
                for (int64_t rp = 0; rp < sched[2 * r + 1]; ++rp) {
                    wsize += orc_on_wolff(s, ndim, dims, n, beta, k0, k1, sid, *sweep, N, work,
                                          work + 2 * N, (int)rp, proj);
                    *sweep += 1;
                }
                continue;
            }
            if (code == 4) {
                for (int64_t rp = 0; rp < sched[2 * r + 1]; ++rp) {
                    orc_on_sw(s, ndim, dims, n, beta, k0, k1, sid, *sweep, N, work, work + N,
                              proj, &c2, &ncl);
                    *sweep += 1;
                }
                continue;
            }
            for (int64_t rp = 0; rp < sched[2 * r + 1]; ++rp) {
                for (int c = 0; c < 2; ++c) {
                    int64_t a = orc_on_half_sweep(s, ndim, dims, n, code, c, beta, step, k0, k1, sid, *sweep);
                    if (a < 0) return -2;
                    acc += a;
                }
                *sweep += 1;
            }
        }
        orc_on_measure(s, ndim, dims, n, &e_out[meas], &m_out[meas * n], &sc_out[meas * ndim],
                       &ss_out[meas * ndim]);
        acc_out[meas] = acc;
        if (aux_out) {
            aux_out[4 * meas + 0] = acc;
            aux_out[4 * meas + 1] = wsize;
            aux_out[4 * meas + 2] = c2;
            aux_out[4 * meas + 3] = ncl;
        }
    }
    free(work);
    free(proj);
    return 0;
}

int orc_on_run(double *s, int ndim, const int32_t *dims, int n, double beta, double step,
               uint32_t k0, uint32_t k1, uint32_t sid, uint64_t *sweep, const int64_t *sched,
               int nrows, int64_t n_meas, double *e_out, double *m_out, double *sc_out,
               double *ss_out, int64_t *acc_out)
{
    return orc_on_run_aux(s, ndim, dims, n, beta, step, k0, k1, sid, sweep, sched, nrows, n_meas,
                          e_out, m_out, sc_out, ss_out, acc_out, NULL);
}
