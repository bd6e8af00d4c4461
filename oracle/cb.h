/* ORACLE (test infrastructure only) -- checkerboard enumeration of a periodic
 * hypercubic lattice (d <= 3, even extents): colour-c sites k = r*M + m, row
 * r = x1 + L1 x2, x0 = 2m + ((x1+x2+c)&1); neighbours in the reference's column
 * order +x,-x,+y,-y,+z,-z (critmc/lattice.py:118-134). */
#ifndef CRITMC_ORACLE_CB_H
#define CRITMC_ORACLE_CB_H
#include <stdint.h>

typedef struct {
    int d, L[3], M;
    int64_t N, half;
} orc_cb;

static inline int orc_cb_init(orc_cb *g, int ndim, const int32_t *dims)
{
    if (ndim < 1 || ndim > 3) return -1;
    g->d = ndim;
    g->N = 1;
    for (int a = 0; a < 3; ++a) g->L[a] = 1;
    for (int a = 0; a < ndim; ++a) {
        if (dims[a] < 2 || (dims[a] & 1)) return -2;
        g->L[a] = dims[a];
        g->N *= dims[a];
    }
    g->M = g->L[0] / 2;
    g->half = g->N / 2;
    return 0;
}

static inline int64_t orc_cb_site(const orc_cb *g, int64_t k, int c, int64_t *nb)
{
    int64_t r = k / g->M;
    int m = (int)(k - r * g->M);
    int x1 = (int)(r % g->L[1]), x2 = (int)(r / g->L[1]);
    int x0 = 2 * m + ((x1 + x2 + c) & 1);
    int64_t base = r * g->L[0];
    nb[0] = base + (x0 + 1 == g->L[0] ? 0 : x0 + 1);
    nb[1] = base + (x0 == 0 ? g->L[0] - 1 : x0 - 1);
    if (g->d >= 2) {
        int64_t row0 = (int64_t)x2 * g->L[1];
        nb[2] = (row0 + (x1 + 1 == g->L[1] ? 0 : x1 + 1)) * g->L[0] + x0;
        nb[3] = (row0 + (x1 == 0 ? g->L[1] - 1 : x1 - 1)) * g->L[0] + x0;
    }
    if (g->d == 3) {
        nb[4] = ((int64_t)(x2 + 1 == g->L[2] ? 0 : x2 + 1) * g->L[1] + x1) * g->L[0] + x0;
        nb[5] = ((int64_t)(x2 == 0 ? g->L[2] - 1 : x2 - 1) * g->L[1] + x1) * g->L[0] + x0;
    }
    return base + x0;
}

static inline uint64_t orc_thr_p(double p)
{
    if (!(p < 1.0)) return 1ull << 32;
    if (!(p > 0.0)) return 0;
    return (uint64_t)ceil(p * 4294967296.0);
}
#endif
