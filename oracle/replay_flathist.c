/*
 * ORACLE (test infrastructure only) -- CPU replay of the flat-histogram walks
 * (Wang-Landau with the 1/t refinement, multicanonical production), the
 * bit-exact target of the batched GPU walkers.
 *
 * Restated from the reference: critmc/kernels/flathist.py:22-132 (moves,
 * window rejection, ln g / histogram updates, recorded series).  Levels:
 * Ising idx = (E_b + B) / 4, Potts idx = E + B.  Move t of a walker draws
 * Philox block 0 at counter t (tag ORC_TAG_FLAT, the walker's stream): word 0
 * -> site floor(x N / 2^32), word 1 -> Potts proposal randbelow(q - 1), +1 if
 * >= old, word 2 -> u = x 2^-32; accepted iff ln_w_new >= ln_w_old or
 * u < det_exp(ln_w_new - ln_w_old) (ln_w = -ln g for Wang-Landau).
 */
#include <stdint.h>

#include "philox.h"

double orc_det_exp(double x);

static int64_t orc_fh_nbr(const int32_t *dims, int ndim, int64_t i, int a, int fwd)
{
    int64_t stride = 1;
    for (int b = 0; b < a && b < ndim; ++b) stride *= dims[b];
    const int64_t L = dims[a], x = (i / stride) % L;
    if (fwd) return x + 1 == L ? i - (L - 1) * stride : i + stride;
    return x == 0 ? i + (L - 1) * stride : i - stride;
}

static int orc_fh_delta(const int8_t *s, int ndim, const int32_t *dims, int q, int64_t i,
                        uint32_t xprop, int *d)
{
    const int si = s[i];
    if (q == 0) {
        int field = 0;
        for (int c = 0; c < 2 * ndim; ++c) field += s[orc_fh_nbr(dims, ndim, i, c >> 1, (c & 1) == 0)];
        *d = (si * field) / 2;
        return -si;
    }
    int nw = (int)(((uint64_t)xprop * (uint32_t)(q - 1)) >> 32);
    if (nw >= si) ++nw;
    int n_old = 0, n_new = 0;
    for (int c = 0; c < 2 * ndim; ++c) {
        int v = s[orc_fh_nbr(dims, ndim, i, c >> 1, (c & 1) == 0)];
        n_old += v == si;
        n_new += v == nw;
    }
    *d = n_old - n_new;
    return nw;
}

static int orc_fh_accept(double ln_w_old, double ln_w_new, uint32_t xu)
{
    return ln_w_new >= ln_w_old || (double)xu * 0x1p-32 < orc_det_exp(ln_w_new - ln_w_old);
}

int orc_wl_run(int8_t *s, int ndim, const int32_t *dims, int q, int32_t *idx, double *ln_g,
               int64_t *hist, int lo, int hi, uint32_t k0, uint32_t k1, uint32_t sid,
               uint64_t *moves, double ln_f, int mode_1t, double n_levels, int64_t n_moves)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    int cur = *idx;
    uint64_t t = *moves;
    for (int64_t m = 0; m < n_moves; ++m, ++t) {
        if (mode_1t) ln_f = n_levels / ((double)t + 1.0);
        uint32_t w[4];
        orc_draw_block(k0, k1, sid, t, 0, ORC_TAG_FLAT, 0, w);
        int64_t i = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
        int d;
        int nw = orc_fh_delta(s, ndim, dims, q, i, w[1], &d);
        int nidx = cur + d;
        if (lo <= nidx && nidx <= hi && orc_fh_accept(-ln_g[cur], -ln_g[nidx], w[2])) {
            s[i] = (int8_t)nw;
            cur = nidx;
        }
        ln_g[cur] += ln_f;
        hist[cur] += 1;
    }
    *idx = cur;
    *moves = t;
    return 0;
}

int orc_muca_run(int8_t *s, int ndim, const int32_t *dims, int q, int32_t *idx, int32_t *order,
                 const double *ln_w, int64_t *hist, int lo, int hi, uint32_t k0, uint32_t k1,
                 uint32_t sid, uint64_t *moves, int64_t n_meas, int64_t moves_per_meas,
                 int32_t *idx_out, int64_t *ord_out)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    int cur = *idx;
    uint64_t t = *moves;
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        for (int64_t m = 0; m < moves_per_meas; ++m, ++t) {
            uint32_t w[4];
            orc_draw_block(k0, k1, sid, t, 0, ORC_TAG_FLAT, 0, w);
            int64_t i = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
            int d;
            int old = s[i];
            int nw = orc_fh_delta(s, ndim, dims, q, i, w[1], &d);
            int nidx = cur + d;
            if (lo <= nidx && nidx <= hi && orc_fh_accept(ln_w[cur], ln_w[nidx], w[2])) {
                s[i] = (int8_t)nw;
                cur = nidx;
                if (q == 0) {
                    order[0] -= 2 * old;
                } else {
                    order[old] -= 1;
                    order[nw] += 1;
                }
            }
            hist[cur] += 1;
        }
        idx_out[meas] = cur;
        if (q == 0) {
            ord_out[meas] = order[0];
        } else {
            int64_t sq = 0;
            for (int k = 0; k < q; ++k) sq += (int64_t)order[k] * order[k];
            ord_out[meas] = sq;
        }
    }
    *idx = cur;
    *moves = t;
    return 0;
}
