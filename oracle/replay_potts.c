/*
 * ORACLE (test infrastructure only) -- CPU replay of the checkerboard Potts
 * schedule, bit-exact target of the GPU kernels.
 *
 * Rules restated from the reference (critmc/kernels/potts.py):
 *   Metropolis :44-71 (boltz[d] = exp(-beta d), :194-197)
 *   heat bath  :74-103 (weight[n] = exp(beta n), :198), FP64 sums in k order
 *   energy / counts :27-41, run :170-235
 * Visiting order: colour 0 then colour 1; within a colour the sites are
 * k = r*M + m (row r = x1 + L1 x2, x0 = 2m + ((x1+x2+c)&1)).  Philox tag
 * POTTS.  Metropolis: site k uses words 2(k&1), 2(k&1)+1 of block k>>1, the
 * proposal and then the acceptance u.  Heat bath: site k uses word k&3 of block
 * k>>2.  randbelow(n) = floor(x n / 2^32).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "philox.h"

#include "cb.h"

int orc_potts_energy(const int8_t *s, int ndim, const int32_t *dims, int q, int64_t *energy,
                     int64_t *counts)
{
    orc_cb g;
    int rc = orc_cb_init(&g, ndim, dims);
    if (rc) return rc;
    int64_t e = 0, nb[6];
    for (int k = 0; k < q; ++k) counts[k] = 0;
    for (int c = 0; c < 2; ++c)
        for (int64_t k = 0; k < g.half; ++k) {
            int64_t i = orc_cb_site(&g, k, c, nb);
            counts[s[i]] += 1;
            for (int j = 0; j < 2 * g.d; j += 2) e -= s[nb[j]] == s[i];
        }
    *energy = e;
    return 0;
}

int64_t orc_cluster_wolff(int8_t *s, int ndim, const int32_t *dims, int q, uint64_t t_add,
                          uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep,
                          int32_t *queue);
void orc_cluster_sw(int8_t *s, int ndim, const int32_t *dims, int q, uint64_t t_add, uint32_t k0,
                    uint32_t k1, uint32_t stream, uint64_t sweep, int32_t *parent, int32_t *size,
                    int64_t *c2, int64_t *ncl);

/* Schedule interpreter with Wolff (3) and Swendsen-Wang (4) rows
 * (critmc/kernels/potts.py:170-235); aux_out [n_meas][4] = (accepted, Wolff
 * sizes, sum |C|^2 and cluster count of the measurement's last SW update). */
int orc_potts_run_aux(int8_t *s, int ndim, const int32_t *dims, int q, double beta, uint32_t k0,
                      uint32_t k1, uint32_t stream, uint64_t *sweep, const int64_t *sched,
                      int nrows, int64_t n_meas, int64_t *energy, int64_t *counts, int64_t *e_out,
                      int64_t *counts_out, int64_t *acc_out, int64_t *aux_out)
{
    orc_cb g;
    int rc = orc_cb_init(&g, ndim, dims);
    if (rc) return rc;
    const int z = 2 * g.d;
    uint64_t T[8];
    double w[8];
    for (int d = 0; d <= z; ++d) {
        T[d] = orc_thr_p(exp(-beta * d));
        w[d] = exp(beta * d);
    }
    int64_t nbi[6];
    const uint64_t t_add = orc_thr_p(1.0 - exp(-beta));
    int32_t *work = NULL;
    for (int row = 0; row < nrows; ++row) {
        int code = (int)sched[2 * row];
        if (code != 0 && code != 2 && code != 3 && code != 4) return -3;
        if ((code == 3 || code == 4) && !aux_out) return -5;
        if ((code == 3 || code == 4) && !work) work = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)g.N);
    }
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t acc = 0, wsize = 0, c2 = 0, ncl = 0;
        for (int row = 0; row < nrows; ++row) {
            int code = (int)sched[2 * row];
            if (code == 3 || code == 4) {
                for (int64_t rp = 0; rp < sched[2 * row + 1]; ++rp) {
                    if (code == 3)
                        wsize += orc_cluster_wolff(s, ndim, dims, q, t_add, k0, k1, stream, *sweep, work);
                    else
                        orc_cluster_sw(s, ndim, dims, q, t_add, k0, k1, stream, *sweep, work,
                                       work + g.N, &c2, &ncl);
                    *sweep += 1;
                }
                orc_potts_energy(s, ndim, dims, q, energy, counts);
                continue;
            }
            for (int64_t rp = 0; rp < sched[2 * row + 1]; ++rp) {
                for (int c = 0; c < 2; ++c) {
                    uint32_t blk[4];
                    for (int64_t k = 0; k < g.half; ++k) {
                        uint32_t xa, xb = 0;
                        if (code == 0) {
                            if ((k & 1) == 0)
                                orc_draw_block(k0, k1, stream, *sweep, c, ORC_TAG_POTTS,
                                               (uint32_t)(k >> 1), blk);
                            xa = blk[2 * (k & 1)];
                            xb = blk[2 * (k & 1) + 1];
                        } else {
                            if ((k & 3) == 0)
                                orc_draw_block(k0, k1, stream, *sweep, c, ORC_TAG_POTTS,
                                               (uint32_t)(k >> 2), blk);
                            xa = blk[k & 3];
                        }
                        int64_t i = orc_cb_site(&g, k, c, nbi);
                        int old = s[i];
                        int nb[6];
                        for (int j = 0; j < z; ++j) nb[j] = s[nbi[j]];
                        if (code == 0) {
                            int nw = (int)(((uint64_t)xa * (uint32_t)(q - 1)) >> 32);
                            if (nw >= old) ++nw;
                            int n_old = 0, n_new = 0;
                            for (int j = 0; j < z; ++j) {
                                n_old += nb[j] == old;
                                n_new += nb[j] == nw;
                            }
                            int delta = n_old - n_new;
                            if (delta <= 0 || (uint64_t)xb < T[delta]) {
                                s[i] = (int8_t)nw;
                                counts[old] -= 1;
                                counts[nw] += 1;
                                *energy += delta;
                                ++acc;
                            }
                        } else {
                            int nk[128];
                            for (int kk = 0; kk < q; ++kk) nk[kk] = 0;
                            for (int j = 0; j < z; ++j) nk[nb[j]] += 1;
                            double total = 0.0;
                            for (int kk = 0; kk < q; ++kk) total += w[nk[kk]];
                            double r = ((double)xa * 0x1p-32) * total;
                            int nw = q - 1;
                            double cum = 0.0;
                            for (int kk = 0; kk < q; ++kk) {
                                cum += w[nk[kk]];
                                if (r < cum) {
                                    nw = kk;
                                    break;
                                }
                            }
                            if (nw != old) {
                                *energy += nk[old] - nk[nw];
                                counts[old] -= 1;
                                counts[nw] += 1;
                                s[i] = (int8_t)nw;
                                ++acc;
                            }
                        }
                    }
                }
                *sweep += 1;
            }
        }
        e_out[meas] = *energy;
        for (int kk = 0; kk < q; ++kk) counts_out[meas * q + kk] = counts[kk];
        acc_out[meas] = acc;
        if (aux_out) {
            aux_out[4 * meas + 0] = acc;
            aux_out[4 * meas + 1] = wsize;
            aux_out[4 * meas + 2] = c2;
            aux_out[4 * meas + 3] = ncl;
        }
    }
    free(work);
    return 0;
}

int orc_potts_run(int8_t *s, int ndim, const int32_t *dims, int q, double beta, uint32_t k0,
                  uint32_t k1, uint32_t stream, uint64_t *sweep, const int64_t *sched, int nrows,
                  int64_t n_meas, int64_t *energy, int64_t *counts, int64_t *e_out,
                  int64_t *counts_out, int64_t *acc_out)
{
    return orc_potts_run_aux(s, ndim, dims, q, beta, k0, k1, stream, sweep, sched, nrows, n_meas,
                             energy, counts, e_out, counts_out, acc_out, NULL);
}

void orc_potts_init(int8_t *s, int64_t n, int q, uint32_t k0, uint32_t k1, uint32_t stream)
{
    uint32_t w[4];
    for (int64_t i = 0; i < n; ++i) {
        if ((i & 3) == 0) orc_draw_block(k0, k1, stream, 0, 0, ORC_TAG_INIT, (uint32_t)(i >> 2), w);
        s[i] = (int8_t)(((uint64_t)w[i & 3] * (uint32_t)q) >> 32);
    }
}
