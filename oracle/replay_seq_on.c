/*
 * ORACLE (test infrastructure only) -- FP64 replay of O(n) chains (n = 2, 3)
 * in the reference's own site orders on any lattice given by its neighbour
 * table nbr[N][z] in pair form (column 2a = +e_a, 2a+1 = -e_a;
 * critmc/lattice.py:129-134, 173-213), the target of the sequential GPU
 * kernel, which also runs in FP64.
 *
 * Restated from critmc/kernels/on.py:
 *   metropolis (0)         sites 0..N-1 in turn (:112-137, random_order False)
 *   metropolis-random (1)  N attempts at uniformly drawn sites (:112-137)
 *   heatbath (2)           sites 0..N-1 in turn (:186-261)
 *   overrelax (5)          sites 0..N-1 in turn (:140-183)
 *   wolff (3), sw (4)      the embedded Ising model along a random axis
 *                          (:264-345), as replay_on.c with the neighbours read
 *                          from the table: bond b = a N + i joins i and
 *                          nbr[i][2a], a < z/2
 *   per measurement        normalise every spin, then E, M[n], and the
 *                          stiffness sums c_a, s_a over the forward bonds in
 *                          the reference's bond order: axis-major, sites
 *                          ascending (:35-87, :404-413)
 * The local field is summed over the neighbour columns in order (:90-97).
 * Each repetition of any row takes one value of the chain's sweep counter.
 *
 * Uniforms (tag ORC_TAG_SEQ, the chain's key and stream; colour field = the
 * update code, plus attempt << 3 for the XY heat bath's rejection attempts).
 * Box-Muller takes u1 = (x+1)/2^32 and u2 = x/2^32:
 * g0 = sqrt(-2 ln u1) cos(2 pi u2), g1 = ... sin(2 pi u2).
 *   metropolis, XY:          block t: w0 -> site floor(w0 N / 2^32) (random order);
 *                            (w1, w2) -> Gaussians g0, g1; w3 = acceptance
 *   metropolis, Heisenberg:  block 2t: (w0, w1) -> g0, g1, (w2, w3) -> g2;
 *                            block 2t+1: w0 = acceptance, w1 -> site
 *   acceptance:              dE <= 0 or x 2^-32 < det_exp(-beta dE)
 *   heat bath, XY:           attempt a < 32 of site t draws block t at field
 *                            2 | a << 3: w0 -> z = cos(pi w0 2^-32), w1 -> u2 =
 *                            (w1+1) 2^-32, top bit of w2 -> sign (Best & Fisher);
 *                            the last proposal is kept if all 32 reject;
 *                            kappa < 1e-8: phi = pi (2 w0 2^-32 - 1), attempt 0
 *   heat bath, Heisenberg:   block t: r = (w0+1) 2^-32, azimuth 2 pi w1 2^-32;
 *                            |h| = 0: z = 2 w2 2^-32 - 1, azimuth 2 pi w3 2^-32
 *   cluster rows:            tag ORC_TAG_CLUSTER as replay_on.c
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "philox.h"

double orc_det_exp(double x);

enum { SQO_HB_ATTEMPTS = 32 };

static void sqo_block(uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, int field,
                      uint32_t t, uint32_t w[4])
{
    orc_draw_block(k0, k1, sid, sweep, field, ORC_TAG_SEQ, t, w);
}

static void sqo_bm(uint32_t xa, uint32_t xb, double *g0, double *g1)
{
    double u1 = ((double)xa + 1.0) * 0x1p-32, u2 = (double)xb * 0x1p-32;
    double r = sqrt(-2.0 * log(u1));
    *g0 = r * cos(2.0 * M_PI * u2);
    *g1 = r * sin(2.0 * M_PI * u2);
}

static void sqo_field(const double *s, const int32_t *nbr, int z, int n, int64_t i, double *h)
{
    for (int c = 0; c < n; ++c) h[c] = 0.0;
    for (int j = 0; j < z; ++j) {
        const double *sj = s + (int64_t)nbr[i * z + j] * n;
        for (int c = 0; c < n; ++c) h[c] += sj[c];
    }
}

/* One Metropolis sweep (code 0 typewriter, 1 random sites); returns accepted moves. */
static int64_t sqo_metropolis(double *s, int64_t N, int z, const int32_t *nbr, int n, int code,
                              double beta, double step, uint32_t k0, uint32_t k1, uint32_t sid,
                              uint64_t sweep)
{
    int64_t acc = 0;
    for (int64_t t = 0; t < N; ++t) {
        uint32_t w[4], v[4];
        double g[4];
        uint32_t xs, xu;
        if (n == 2) {
            sqo_block(k0, k1, sid, sweep, code, (uint32_t)t, w);
            sqo_bm(w[1], w[2], &g[0], &g[1]);
            xs = w[0];
            xu = w[3];
        } else {
            sqo_block(k0, k1, sid, sweep, code, (uint32_t)(2 * t), w);
            sqo_block(k0, k1, sid, sweep, code, (uint32_t)(2 * t + 1), v);
            sqo_bm(w[0], w[1], &g[0], &g[1]);
            sqo_bm(w[2], w[3], &g[2], &g[3]);
            xu = v[0];
            xs = v[1];
        }
        const int64_t i = code == 1 ? (int64_t)(((uint64_t)xs * (uint64_t)N) >> 32) : t;
        double h[3], tr[3], nrm = 0.0, de = 0.0;
        sqo_field(s, nbr, z, n, i, h);
        double *si = s + i * n;
        for (int c = 0; c < n; ++c) {
            tr[c] = si[c] + step * g[c];
            nrm += tr[c] * tr[c];
        }
        nrm = sqrt(nrm);
        for (int c = 0; c < n; ++c) {
            tr[c] /= nrm;
            de -= (tr[c] - si[c]) * h[c];
        }
        if (de <= 0.0 || (double)xu * 0x1p-32 < orc_det_exp(-beta * de)) {
            for (int c = 0; c < n; ++c) si[c] = tr[c];
            ++acc;
        }
    }
    return acc;
}

/* Best & Fisher von Mises deviate (critmc/kernels/on.py:186-204). */
static double sqo_von_mises(double kappa, uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep,
                            uint32_t t)
{
    uint32_t b[4];
    if (kappa < 1e-8) {
        sqo_block(k0, k1, sid, sweep, 2, t, b);
        return M_PI * (2.0 * ((double)b[0] * 0x1p-32) - 1.0);
    }
    double root = sqrt(1.0 + 4.0 * kappa * kappa);
    double tau = 1.0 + root;
    double rho = 2.0 * kappa * tau / ((root + 1.0) * (tau + sqrt(2.0 * tau)));
    double r = (1.0 + rho * rho) / (2.0 * rho);
    double f = 0.0;
    uint32_t sign = 0;
    for (int a = 0; a < SQO_HB_ATTEMPTS; ++a) {
        sqo_block(k0, k1, sid, sweep, 2 | (a << 3), t, b);
        double zz = cos(M_PI * ((double)b[0] * 0x1p-32));
        f = (1.0 + r * zz) / (r + zz);
        double c = kappa * (r - f);
        double u2 = ((double)b[1] + 1.0) * 0x1p-32;
        sign = b[2] >> 31;
        if (c * (2.0 - c) - u2 > 0.0 || log(c / u2) + 1.0 - c >= 0.0) break;
    }
    double theta = acos(fmin(1.0, fmax(-1.0, f)));
    return sign == 0 ? theta : -theta;
}

static void sqo_heatbath(double *s, int64_t N, int z, const int32_t *nbr, int n, double beta,
                         uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep)
{
    for (int64_t t = 0; t < N; ++t) {
        double h[3];
        sqo_field(s, nbr, z, n, t, h);
        double *si = s + t * n;
        double hn2 = 0.0;
        for (int c = 0; c < n; ++c) hn2 += h[c] * h[c];
        double hn = sqrt(hn2);
        double kappa = beta * hn;
        if (n == 2) {
            double phi0 = hn > 0.0 ? atan2(h[1], h[0]) : 0.0;
            double phi = phi0 + sqo_von_mises(kappa, k0, k1, sid, sweep, (uint32_t)t);
            si[0] = cos(phi);
            si[1] = sin(phi);
            continue;
        }
        uint32_t b[4];
        sqo_block(k0, k1, sid, sweep, 2, (uint32_t)t, b);
        if (hn == 0.0) {
            double zz = 2.0 * ((double)b[2] * 0x1p-32) - 1.0, az = 2.0 * M_PI * ((double)b[3] * 0x1p-32);
            double w = sqrt(fmax(0.0, 1.0 - zz * zz));
            si[0] = w * cos(az);
            si[1] = w * sin(az);
            si[2] = zz;
            continue;
        }
        double r = ((double)b[0] + 1.0) * 0x1p-32;
        double u = kappa < 1e-8 ? 2.0 * r - 1.0 : 1.0 + log(r + (1.0 - r) * exp(-2.0 * kappa)) / kappa;
        u = fmin(1.0, fmax(-1.0, u));
        double hx = h[0] / hn, hy = h[1] / hn, hz = h[2] / hn;
        double ax = fabs(hx) < 0.9 ? 1.0 : 0.0, ay = 1.0 - ax;
        double proj = ax * hx + ay * hy;
        double e1[3] = {ax - proj * hx, ay - proj * hy, -proj * hz};
        double nrm = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (int c = 0; c < 3; ++c) e1[c] /= nrm;
        double e2[3] = {hy * e1[2] - hz * e1[1], hz * e1[0] - hx * e1[2], hx * e1[1] - hy * e1[0]};
        double phi = 2.0 * M_PI * ((double)b[1] * 0x1p-32);
        double w = sqrt(fmax(0.0, 1.0 - u * u));
        double cp = cos(phi), sp = sin(phi);
        si[0] = u * hx + w * (cp * e1[0] + sp * e2[0]);
        si[1] = u * hy + w * (cp * e1[1] + sp * e2[1]);
        si[2] = u * hz + w * (cp * e1[2] + sp * e2[2]);
    }
}

static void sqo_overrelax(double *s, int64_t N, int z, const int32_t *nbr, int n)
{
    for (int64_t t = 0; t < N; ++t) {
        double h[3];
        sqo_field(s, nbr, z, n, t, h);
        double *si = s + t * n;
        double hh = 0.0, sh = 0.0;
        for (int c = 0; c < n; ++c) {
            hh += h[c] * h[c];
            sh += si[c] * h[c];
        }
        if (hh == 0.0) continue;
        double f = 2.0 * sh / hh;
        for (int c = 0; c < n; ++c) si[c] = f * h[c] - si[c];
    }
}

static uint32_t sqo_cl_word(uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, int field,
                            int64_t index)
{
    uint32_t w[4];
    orc_draw_block(k0, k1, sid, sweep, field, ORC_TAG_CLUSTER, (uint32_t)(index >> 2), w);
    return w[index & 3];
}

static void sqo_axis(int n, uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, double *r)
{
    uint32_t w[4];
    orc_draw_block(k0, k1, sid, sweep, 3, ORC_TAG_CLUSTER, 0, w);
    if (n == 2) {
        double phi = 2.0 * M_PI * ((double)w[0] * 0x1p-32);
        r[0] = cos(phi);
        r[1] = sin(phi);
    } else {
        double zz = 2.0 * ((double)w[0] * 0x1p-32) - 1.0, az = 2.0 * M_PI * ((double)w[1] * 0x1p-32);
        double ww = sqrt(fmax(0.0, 1.0 - zz * zz));
        r[0] = ww * cos(az);
        r[1] = ww * sin(az);
        r[2] = zz;
    }
}

static double sqo_proj(const double *s, int n, int64_t i, const double *r)
{
    double p = 0.0;
    for (int c = 0; c < n; ++c) p += s[i * n + c] * r[c];
    return p;
}

static int sqo_active(double x, double beta, uint32_t w)
{
    return x > 0.0 && (double)w * 0x1p-32 < 1.0 - orc_det_exp(-2.0 * beta * x);
}

static int32_t sqo_find(int32_t *parent, int32_t i)
{
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
}

static void sqo_sw(double *s, int64_t N, int z, const int32_t *nbr, int n, double beta, uint32_t k0,
                   uint32_t k1, uint32_t sid, uint64_t sweep, int32_t *parent, int32_t *size,
                   double *proj, int64_t *c2, int64_t *ncl)
{
    double r[3];
    sqo_axis(n, k0, k1, sid, sweep, r);
    for (int64_t i = 0; i < N; ++i) {
        parent[i] = (int32_t)i;
        proj[i] = sqo_proj(s, n, i, r);
    }
    for (int a = 0; a < z / 2; ++a)
        for (int64_t i = 0; i < N; ++i) {
            int64_t j = nbr[i * z + 2 * a];
            if (!sqo_active(proj[i] * proj[j], beta, sqo_cl_word(k0, k1, sid, sweep, 0, (int64_t)a * N + i)))
                continue;
            int32_t ri = sqo_find(parent, (int32_t)i), rj = sqo_find(parent, (int32_t)j);
            if (ri == rj) continue;
            if (ri < rj) parent[rj] = ri;
            else parent[ri] = rj;
        }
    for (int64_t i = 0; i < N; ++i) size[i] = 0;
    for (int64_t i = 0; i < N; ++i) size[sqo_find(parent, (int32_t)i)] += 1;
    *c2 = 0;
    *ncl = 0;
    for (int64_t i = 0; i < N; ++i) {
        int32_t rt = sqo_find(parent, (int32_t)i);
        if (rt == i) {
            *c2 += (int64_t)size[i] * size[i];
            *ncl += 1;
        }
        if ((sqo_cl_word(k0, k1, sid, sweep, 1, rt) >> 31) == 0)
            for (int c = 0; c < n; ++c) s[i * n + c] -= 2.0 * proj[i] * r[c];
    }
}

static int64_t sqo_wolff(double *s, int64_t N, int z, const int32_t *nbr, int n, double beta,
                         uint32_t k0, uint32_t k1, uint32_t sid, uint64_t sweep, int32_t *queue,
                         int32_t *mark, int tag, double *proj)
{
    double r[3];
    sqo_axis(n, k0, k1, sid, sweep, r);
    int64_t seed = (int64_t)(((uint64_t)sqo_cl_word(k0, k1, sid, sweep, 2, 0) * (uint64_t)N) >> 32);
    double p0 = sqo_proj(s, n, seed, r);
    proj[seed] = p0;
    for (int c = 0; c < n; ++c) s[seed * n + c] -= 2.0 * p0 * r[c];
    mark[seed] = tag;
    queue[0] = (int32_t)seed;
    int64_t head = 0, tail = 1;
    while (head < tail) {
        int64_t k = queue[head++];
        double pk = proj[k];
        for (int cc = 0; cc < z; ++cc) {
            int64_t j = nbr[k * z + cc];
            if (mark[j] == tag) continue;
            double pj = sqo_proj(s, n, j, r);
            int64_t b = (int64_t)(cc >> 1) * N + ((cc & 1) ? j : k);
            if (!sqo_active(pk * pj, beta, sqo_cl_word(k0, k1, sid, sweep, 0, b))) continue;
            proj[j] = pj;
            for (int c = 0; c < n; ++c) s[j * n + c] -= 2.0 * pj * r[c];
            mark[j] = tag;
            queue[tail++] = (int32_t)j;
        }
    }
    return tail;
}

/* Normalise every spin, then E, M[n], c_a, s_a over the forward bonds,
 * axis-major (critmc/kernels/on.py:35-87). */
static void sqo_measure(double *s, int64_t N, int z, const int32_t *nbr, int n, double *e,
                        double *m, double *sc, double *ss)
{
    for (int64_t i = 0; i < N; ++i) {
        double n2 = 0.0;
        for (int c = 0; c < n; ++c) n2 += s[i * n + c] * s[i * n + c];
        double nrm = sqrt(n2);
        for (int c = 0; c < n; ++c) s[i * n + c] /= nrm;
    }
    *e = 0.0;
    for (int a = 0; a < z / 2; ++a)
        for (int64_t i = 0; i < N; ++i) {
            int64_t j = nbr[i * z + 2 * a];
            double dot = 0.0;
            for (int c = 0; c < n; ++c) dot += s[i * n + c] * s[j * n + c];
            *e -= dot;
        }
    for (int c = 0; c < n; ++c) m[c] = 0.0;
    for (int64_t i = 0; i < N; ++i)
        for (int c = 0; c < n; ++c) m[c] += s[i * n + c];
    for (int a = 0; a < z / 2; ++a) {
        sc[a] = 0.0;
        ss[a] = 0.0;
        for (int64_t i = 0; i < N; ++i) {
            int64_t j = nbr[i * z + 2 * a];
            sc[a] += s[i * n] * s[j * n] + s[i * n + 1] * s[j * n + 1];
            ss[a] += s[i * n] * s[j * n + 1] - s[i * n + 1] * s[j * n];
        }
    }
}

/* Apply the schedule n_meas times; outputs E [n_meas], M [n_meas][n], c / s
 * [n_meas][z/2] and aux [n_meas][4] = (accepted, Wolff sizes, sum |C|^2, clusters). */
int orc_seq_on_run(double *s, int64_t N, int z, const int32_t *nbr, int n, double beta,
                   double step, uint32_t k0, uint32_t k1, uint32_t sid, uint64_t *sweep,
                   const int64_t *sched, int nrows, int64_t n_meas, double *e_out, double *m_out,
                   double *sc_out, double *ss_out, int64_t *aux_out)
{
    if (n != 2 && n != 3) return -1;
    int32_t *work = NULL;
    double *proj = NULL;
    for (int r = 0; r < nrows; ++r) {
        int code = (int)sched[2 * r];
        if (code < 0 || code > 5) return -3;
        if ((code == 3 || code == 4) && !work) {
            work = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)N);
            proj = (double *)malloc(sizeof(double) * (size_t)N);
        }
    }
    const int A = z / 2;
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t acc = 0, wsize = 0, c2 = 0, ncl = 0;
        for (int r = 0; r < nrows; ++r) {
            const int code = (int)sched[2 * r];
            if (code == 3)
                for (int64_t i = 0; i < N; ++i) work[2 * N + i] = -1;
            for (int64_t rp = 0; rp < sched[2 * r + 1]; ++rp) {
                if (code == 0 || code == 1)
                    acc += sqo_metropolis(s, N, z, nbr, n, code, beta, step, k0, k1, sid, *sweep);
                else if (code == 2)
                    sqo_heatbath(s, N, z, nbr, n, beta, k0, k1, sid, *sweep);
                else if (code == 5)
                    sqo_overrelax(s, N, z, nbr, n);
                else if (code == 3)
                    wsize += sqo_wolff(s, N, z, nbr, n, beta, k0, k1, sid, *sweep, work, work + 2 * N,
                                       (int)rp, proj);
                else
                    sqo_sw(s, N, z, nbr, n, beta, k0, k1, sid, *sweep, work, work + N, proj, &c2, &ncl);
                *sweep += 1;
            }
        }
        sqo_measure(s, N, z, nbr, n, &e_out[meas], &m_out[meas * n], &sc_out[meas * A],
                    &ss_out[meas * A]);
        aux_out[4 * meas + 0] = acc;
        aux_out[4 * meas + 1] = wsize;
        aux_out[4 * meas + 2] = c2;
        aux_out[4 * meas + 3] = ncl;
    }
    free(work);
    free(proj);
    return 0;
}
