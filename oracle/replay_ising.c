/*
 * ORACLE (test infrastructure only) -- CPU replay of the checkerboard Ising
 * schedule.  Only tests/, the package smoke check and bench.py's CPU-baseline
 * leg load this code, and only as the checker; the product never links it.
 *
 * Acceptance rules restated from the reference:
 *   Metropolis  critmc/kernels/ising.py:47-56 (table), :68-87 (sweep)
 *   heat bath   critmc/kernels/ising.py:59-65 (table), :90-107 (sweep)
 *   E_b, M      critmc/kernels/ising.py:31-44
 *   run loop    critmc/kernels/ising.py:173-237 (records E_b, M, accepted)
 * Geometry: periodic hypercubic, x-fastest site numbering
 *   (critmc/lattice.py:118-134).
 *
 * What differs from the reference on purpose (a different valid Markov chain
 * with the same stationary law): sites are visited colour 0 then colour 1
 * (colour = sum of coordinates mod 2) instead of typewriter order, and the
 * uniform u of a site is the 32-bit number u = x / 2^32 described below,
 * compared as "x < ceil(p * 2^32)", which is "u < p" exactly.
 *
 * Uniform of a site (the stream contract shared with the GPU kernels):
 *   the colour-c sites of a lattice row r (all coordinates but x0 fixed) are
 *   x0 = 2m + ((x1 + x2 + ... + c) & 1), m = 0 .. L0/2-1.  They form groups of
 *   32 consecutive m: group g = r * ceil(L0/64) + m / 32, bit b = m % 32.
 *   Group g draws the 32 words W_0..W_31, W_j = word (j & 3) of the Philox
 *   block with counter (8g + (j >> 2), sweep_lo, sweep_hi | c << 16 |
 *   TAG_ISING << 24, stream).  The site's x has bit (31 - j) equal to bit b
 *   of W_j: the words are bit-planes, most significant plane first.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "philox.h"

#define ORC_METROPOLIS 0
#define ORC_HEATBATH 2
#define ORC_WOLFF 3
#define ORC_SW 4

typedef struct {
    int d;
    int L[4];
    int64_t stride[4];
    int64_t N;
    int64_t rows;
    int M;
    int G;
    int32_t *nbr; /* [N][2d] neighbour table, columns +e_a, -e_a (lattice.py:129-134) */
} orc_geom;

static int orc_geom_init(orc_geom *g, int ndim, const int32_t *dims)
{
    if (ndim < 1 || ndim > 4) return -1;
    g->d = ndim;
    g->N = 1;
    for (int a = 0; a < ndim; ++a) {
        if (dims[a] < 2 || (dims[a] & 1)) return -2; /* checkerboard needs even extents */
        g->L[a] = dims[a];
        g->stride[a] = g->N;
        g->N *= dims[a];
    }
    g->rows = g->N / g->L[0];
    g->M = g->L[0] / 2;
    g->G = (g->M + 31) / 32;
    g->nbr = NULL;
    return 0;
}

static inline int64_t orc_nbr_slow(const orc_geom *g, int64_t i, int a, int dir)
{
    int64_t xa = (i / g->stride[a]) % g->L[a];
    int64_t xn = xa + dir;
    if (xn < 0) xn += g->L[a];
    else if (xn >= g->L[a]) xn -= g->L[a];
    return i + (xn - xa) * g->stride[a];
}

static int orc_geom_table(orc_geom *g)
{
    const int z = 2 * g->d;
    g->nbr = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->N * z);
    if (!g->nbr) return -4;
    for (int64_t i = 0; i < g->N; ++i)
        for (int a = 0; a < g->d; ++a) {
            g->nbr[i * z + 2 * a] = (int32_t)orc_nbr_slow(g, i, a, +1);
            g->nbr[i * z + 2 * a + 1] = (int32_t)orc_nbr_slow(g, i, a, -1);
        }
    return 0;
}

static inline int64_t orc_nbr(const orc_geom *g, int64_t i, int a, int dir)
{
    if (g->nbr) return g->nbr[i * 2 * g->d + 2 * a + (dir < 0)];
    return orc_nbr_slow(g, i, a, dir);
}

static inline int orc_row_parity(const orc_geom *g, int64_t r)
{
    int64_t i = r * g->L[0];
    int p = 0;
    for (int a = 1; a < g->d; ++a) p += (int)((i / g->stride[a]) % g->L[a]);
    return p & 1;
}

static inline uint64_t orc_threshold(double p)
{
    /* "u < p" for u = x / 2^32 is "x < ceil(p 2^32)"; p >= 1 accepts always */
    if (!(p < 1.0)) return 1ull << 32;
    if (!(p > 0.0)) return 0;
    return (uint64_t)ceil(p * 4294967296.0);
}

/* critmc/kernels/ising.py:47-65: Metropolis table [(a+z)*2 + t], heat-bath table [f+z]. */
void orc_ising_thresholds(double beta, double h, int z, uint64_t *met, uint64_t *hb)
{
    for (int a = -z; a <= z; ++a) {
        for (int t = 0; t < 2; ++t) {
            int spin = 2 * t - 1;
            double d_e = 2.0 * a + 2.0 * h * spin;
            double p = d_e <= 0.0 ? 1.0 : exp(-beta * d_e);
            met[(a + z) * 2 + t] = orc_threshold(p);
        }
        hb[a + z] = orc_threshold(1.0 / (1.0 + exp(-2.0 * beta * (a + h))));
    }
}

void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    orc_philox4x32_10(ctr, key, out);
}

/* Bit-plane words of one group; the site at bit b has x = sum_j bit_b(W_j) 2^(31-j). */
static void orc_group_words(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep, int color,
                            uint32_t group, uint32_t W[32])
{
    for (uint32_t q = 0; q < 8; ++q)
        orc_draw_block(k0, k1, stream, sweep, color, ORC_TAG_ISING, 8u * group + q, &W[4 * q]);
}

uint32_t orc_ising_site_uniform(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep,
                                int color, uint32_t group, int bit)
{
    uint32_t W[32], x = 0;
    orc_group_words(k0, k1, stream, sweep, color, group, W);
    for (int j = 0; j < 32; ++j) x = (x << 1) | ((W[j] >> bit) & 1u);
    return x;
}

static void orc_half_sweep(const orc_geom *g, int8_t *s, int code, int color, const uint64_t *met,
                           const uint64_t *hb, uint32_t k0, uint32_t k1, uint32_t stream,
                           uint64_t sweep, int64_t *e_b, int64_t *m, int64_t *acc)
{
    const int z = 2 * g->d;
    uint32_t W[32];
    for (int64_t r = 0; r < g->rows; ++r) {
        int p = (orc_row_parity(g, r) + color) & 1;
        for (int w = 0; w < g->G; ++w) {
            uint32_t group = (uint32_t)(r * g->G + w);
            orc_group_words(k0, k1, stream, sweep, color, group, W);
            int nb = g->M - 32 * w;
            if (nb > 32) nb = 32;
            for (int b = 0; b < nb; ++b) {
                uint32_t x = 0;
                for (int j = 0; j < 32; ++j) x = (x << 1) | ((W[j] >> b) & 1u);
                int64_t i = r * g->L[0] + 2 * (32 * w + b) + p;
                int field = 0;
                for (int a = 0; a < g->d; ++a)
                    field += s[orc_nbr(g, i, a, +1)] + s[orc_nbr(g, i, a, -1)];
                int si = s[i];
                if (code == ORC_METROPOLIS) {
                    int aa = si * field;
                    uint64_t T = met[(aa + z) * 2 + ((si + 1) >> 1)];
                    if ((uint64_t)x < T) {
                        s[i] = (int8_t)(-si);
                        *e_b += 2 * aa;
                        *m -= 2 * si;
                        *acc += 1;
                    }
                } else {
                    int nw = ((uint64_t)x < hb[field + z]) ? 1 : -1;
                    if (nw != si) {
                        *e_b += 2 * si * field;
                        *m += 2 * nw;
                        s[i] = (int8_t)nw;
                        *acc += 1;
                    }
                }
            }
        }
    }
}

/* Exact integers E_b = -sum_<ij> s_i s_j and M = sum_i s_i (ising.py:31-44). */
int orc_ising_energy(const int8_t *s, int ndim, const int32_t *dims, int64_t *e_b, int64_t *m)
{
    orc_geom g;
    int rc = orc_geom_init(&g, ndim, dims);
    if (rc) return rc;
    int64_t e = 0, mm = 0;
    for (int64_t i = 0; i < g.N; ++i) {
        mm += s[i];
        for (int a = 0; a < g.d; ++a) e -= s[i] * s[orc_nbr(&g, i, a, +1)];
    }
    *e_b = e;
    *m = mm;
    return 0;
}

/*
 * Execute `schedule` (rows of (code, count)) n_meas times; record E_b, M and
 * the accepted count after each application (ising.py:173-237).  `sweep` is
 * the chain's local-sweep counter, advanced by one per Metropolis or
 * heat-bath sweep.  Returns 0, or <0 for an unsupported code or geometry.
 */
/* Wolff / Swendsen-Wang (replay_cluster.c, q = 0 selects the Ising rules). */
int64_t orc_cluster_wolff(int8_t *s, int ndim, const int32_t *dims, int q, uint64_t t_add,
                          uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep,
                          int32_t *queue);
void orc_cluster_sw(int8_t *s, int ndim, const int32_t *dims, int q, uint64_t t_add, uint32_t k0,
                    uint32_t k1, uint32_t stream, uint64_t sweep, int32_t *parent, int32_t *size,
                    int64_t *c2, int64_t *ncl);

int orc_ising_run_aux(int8_t *s, int ndim, const int32_t *dims, double beta, double h,
                      uint32_t k0, uint32_t k1, uint32_t stream, uint64_t *sweep,
                      const int64_t *sched, int nrows, int64_t n_meas, int64_t *e_b, int64_t *m,
                      int64_t *e_out, int64_t *m_out, int64_t *acc_out, int64_t *aux_out)
{
    orc_geom g;
    int rc = orc_geom_init(&g, ndim, dims);
    if (rc) return rc;
    if ((rc = orc_geom_table(&g))) return rc;
    const int z = 2 * g.d;
    uint64_t met[2 * 9 * 2], hb[2 * 9];
    orc_ising_thresholds(beta, h, z, met, hb);
    const uint64_t t_add = orc_threshold(1.0 - exp(-2.0 * beta));
    int32_t *work = NULL;
    for (int row = 0; row < nrows; ++row) {
        int64_t code = sched[2 * row];
        if (code != ORC_METROPOLIS && code != ORC_HEATBATH && code != ORC_WOLFF && code != ORC_SW) {
            free(g.nbr);
            return -3;
        }
        if ((code == ORC_WOLFF || code == ORC_SW) && (h != 0.0 || !aux_out)) {
            free(g.nbr);
            return -5;
        }
        if ((code == ORC_WOLFF || code == ORC_SW) && !work)
            work = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)g.N);
    }
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t acc = 0, wsize = 0, c2 = 0, ncl = 0;
        for (int row = 0; row < nrows; ++row) {
            int64_t code = sched[2 * row], count = sched[2 * row + 1];
            for (int64_t k = 0; k < count; ++k) {
                if (code == ORC_WOLFF) {
                    wsize += orc_cluster_wolff(s, ndim, dims, 0, t_add, k0, k1, stream, *sweep, work);
                } else if (code == ORC_SW) {
                    orc_cluster_sw(s, ndim, dims, 0, t_add, k0, k1, stream, *sweep, work,
                                   work + g.N, &c2, &ncl);
                } else {
                    orc_half_sweep(&g, s, (int)code, 0, met, hb, k0, k1, stream, *sweep, e_b, m, &acc);
                    orc_half_sweep(&g, s, (int)code, 1, met, hb, k0, k1, stream, *sweep, e_b, m, &acc);
                }
                *sweep += 1;
            }
            if (code == ORC_WOLFF || code == ORC_SW) orc_ising_energy(s, ndim, dims, e_b, m);
        }
        e_out[meas] = *e_b;
        m_out[meas] = *m;
        acc_out[meas] = acc;
        if (aux_out) {
            aux_out[4 * meas + 0] = acc;
            aux_out[4 * meas + 1] = wsize;
            aux_out[4 * meas + 2] = c2;
            aux_out[4 * meas + 3] = ncl;
        }
    }
    free(work);
    free(g.nbr);
    return 0;
}

int orc_ising_run(int8_t *s, int ndim, const int32_t *dims, double beta, double h, uint32_t k0,
                  uint32_t k1, uint32_t stream, uint64_t *sweep, const int64_t *sched, int nrows,
                  int64_t n_meas, int64_t *e_b, int64_t *m, int64_t *e_out, int64_t *m_out,
                  int64_t *acc_out)
{
    return orc_ising_run_aux(s, ndim, dims, beta, h, k0, k1, stream, sweep, sched, nrows, n_meas,
                             e_b, m, e_out, m_out, acc_out, NULL);
}

/* Random start: site i takes bit 31 of word (i & 3) of block i >> 2 (sweep 0, colour 0). */
void orc_init_spins(int8_t *s, int64_t n, uint32_t k0, uint32_t k1, uint32_t stream)
{
    uint32_t w[4];
    for (int64_t i = 0; i < n; ++i) {
        if ((i & 3) == 0) orc_draw_block(k0, k1, stream, 0, 0, ORC_TAG_INIT, (uint32_t)(i >> 2), w);
        s[i] = (w[i & 3] >> 31) ? 1 : -1;
    }
}

/* ---- CPU baselines (bench.py's cpu_baseline leg) ------------------------ */

/* Many independent chains of the checkerboard replay, OpenMP over chains. */
int orc_ising_run_batch(int8_t *s, int R, int ndim, const int32_t *dims, const double *betas,
                        double h, const uint32_t *keys, const uint32_t *sids, uint64_t *sweeps,
                        const int64_t *sched, int nrows, int64_t n_meas, int64_t *e_b,
                        int64_t *m, int64_t *e_out, int64_t *m_out, int64_t *acc_out, int threads)
{
    orc_geom g;
    int rc = orc_geom_init(&g, ndim, dims);
    if (rc) return rc;
    int bad = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic) reduction(| : bad)
    for (int r = 0; r < R; ++r) {
        bad |= orc_ising_run(s + (int64_t)r * g.N, ndim, dims, betas[r], h, keys[2 * r],
                             keys[2 * r + 1], sids[r], &sweeps[r], sched, nrows, n_meas, &e_b[r],
                             &m[r], e_out + (int64_t)r * n_meas, m_out + (int64_t)r * n_meas,
                             acc_out + (int64_t)r * n_meas) != 0;
    }
    return bad ? -3 : 0;
}

/*
 * The reference's own sweep order: typewriter Metropolis
 * (critmc/kernels/ising.py:68-87) -- sites 0..N-1 in turn, a uniform drawn
 * only when p < 1 (the `p >= 1.0 or rng.random() < p` short circuit).  Draws
 * come from the chain's Philox stream consumed sequentially, four per block.
 * A timing baseline with the reference's cost structure; its chain is not
 * the checkerboard chain.
 */
int orc_ising_typewriter_batch(int8_t *s, int R, int ndim, const int32_t *dims,
                               const double *betas, double h, const uint32_t *keys,
                               int64_t n_sweeps, int64_t *e_b, int64_t *m, int threads)
{
    orc_geom g;
    int rc = orc_geom_init(&g, ndim, dims);
    if (rc) return rc;
    if ((rc = orc_geom_table(&g))) return rc;
    const int z = 2 * g.d;
#pragma omp parallel for num_threads(threads) schedule(dynamic)
    for (int r = 0; r < R; ++r) {
        int8_t *sr = s + (int64_t)r * g.N;
        uint64_t met[2 * 9 * 2], hb[2 * 9];
        orc_ising_thresholds(betas[r], h, z, met, hb);
        uint32_t buf[4];
        uint32_t blk = 0;
        int used = 4;
        int64_t eb = e_b[r], mm = m[r];
        for (int64_t sw = 0; sw < n_sweeps; ++sw) {
            for (int64_t i = 0; i < g.N; ++i) {
                int field = 0;
                for (int a = 0; a < g.d; ++a)
                    field += sr[orc_nbr(&g, i, a, +1)] + sr[orc_nbr(&g, i, a, -1)];
                int si = sr[i];
                int aa = si * field;
                uint64_t T = met[(aa + z) * 2 + ((si + 1) >> 1)];
                int accept = T >> 32 ? 1 : 0;
                if (!accept) {
                    if (used == 4) {
                        orc_draw_block(keys[2 * r], keys[2 * r + 1], (uint32_t)r, 0, 0, ORC_TAG_TYPEWRITER, blk++,
                                       buf);
                        used = 0;
                    }
                    accept = (uint64_t)buf[used++] < T;
                }
                if (accept) {
                    sr[i] = (int8_t)(-si);
                    eb += 2 * aa;
                    mm -= 2 * si;
                }
            }
        }
        e_b[r] = eb;
        m[r] = mm;
    }
    free(g.nbr);
    return 0;
}
