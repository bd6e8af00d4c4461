/*
 * ORACLE (test infrastructure only) -- CPU replay of batched
 * Edwards-Anderson parallel tempering, bit-exact target of the GPU kernels
 * for spins, permutations, labels, trips, swap counts and the integer q
 * series; with +-J couplings (jtype 0) also for the int32 energies and E
 * series.  With FP64 couplings (jtype 1, Gaussian disorder) the field is
 * h = sum_k J_ik s_k added in column order (+x, -x, +y, ...), Metropolis
 * accepts iff dE <= 0 or u < det_exp(-beta dE), the heat bath sets +1 iff
 * u < 1/(1 + det_exp(-2 beta h)) (u = x 2^-32), and energies are recomputed at
 * every bin edge after the first (critmc/kernels/glass.py:116-122); the FP64
 * energies then agree with the device to rounding (its sums run in another
 * order), and the FP64 accumulators always do.
 *
 * Restated from the reference (critmc/kernels/glass.py):
 *   energy :33-38; Metropolis sweep :41-54; heat bath = Ising heat-bath rule
 *   (critmc/kernels/ising.py:59-65) with local field h = sum J s;
 *   swaps :129-137; labels / trips / label_hist :139-152;
 *   _measure :56-77 and the binned accumulators :153-163 (columns :23-30).
 * Step order per PT step: every configuration of both sets sweeps
 * `sweeps_per_step` times, then set 0 swaps, then set 1 swaps, then every
 * temperature is measured.  Bin of the call's step t (from 0): floor(log2(t + 1));
 * random numbers are keyed by the global step step0 + t.
 *
 * Uniforms.  Per-configuration path (Gaussian couplings, more than 32
 * temperatures, or lattices the multi-spin kernels do not take): configuration
 * c of set r, colour site k uses word k&3 of Philox block k>>2 at (its own
 * sweep counter, colour, tag EA, stream r*m + c).  Multi-spin path
 * (`multispin`, +-J couplings, m <= 32, one bit per configuration): the set's 32 bit planes
 * for colour site k are words 0..3 of Philox blocks 8k .. 8k+7 at (the set's
 * sweep counter = that of its configuration 0, colour, tag EA, stream r), and
 * configuration c's uniform is x = sum_t bit_c(plane_t) 2^(31-t); every
 * configuration of the set then carries the set's counter.
 *
 * det_exp(x): e^x by Cody-Waite reduction x = n ln2 + r and the degree-13
 * Taylor polynomial of e^r in Horner/fma form -- a fixed sequence of IEEE
 * operations, so any correct implementation of it gives the same bits.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "cb.h"
#include "philox.h"

double orc_det_exp(double x)
{
    if (x < -745.0) return 0.0;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    double n = rint(x * inv_ln2);
    double r = fma(-n, ln2_hi, x);
    r = fma(-n, ln2_lo, r);
    static const double c[14] = {1.0,          1.0,           0.5,           1.0 / 6.0,
                                 1.0 / 24.0,   1.0 / 120.0,   1.0 / 720.0,   1.0 / 5040.0,
                                 1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0,
                                 1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0};
    double p = c[13];
    for (int k = 12; k >= 0; --k) p = fma(p, r, c[k]);
    return ldexp(p, (int)n);
}

static int64_t ea_nbr_fwd(int64_t i, int ax, int L)
{
    int64_t stride = ax == 0 ? 1 : (ax == 1 ? L : (int64_t)L * L);
    int x = (int)((i / stride) % L);
    return i + (x + 1 == L ? -(int64_t)(L - 1) * stride : stride);
}

double orc_ea_energy_f64(const int8_t *s, const double *J, int ndim, int L)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= L;
    double e = 0.0;
    for (int64_t i = 0; i < N; ++i)
        for (int ax = 0; ax < ndim; ++ax) e -= J[ax * N + i] * (double)(s[i] * s[ea_nbr_fwd(i, ax, L)]);
    return e;
}

int32_t orc_ea_energy(const int8_t *s, const int8_t *J, int ndim, int L)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= L;
    int32_t e = 0;
    for (int64_t i = 0; i < N; ++i)
        for (int ax = 0; ax < ndim; ++ax) e -= J[ax * N + i] * s[i] * s[ea_nbr_fwd(i, ax, L)];
    return e;
}

int orc_ea_pt_run(int8_t *spins, const void *Jv, int jtype, int ndim, int L, int m,
                  const double *betas, int code, int sweeps_per_step, uint32_t k0, uint32_t k1,
                  uint64_t *sweeps, void *energies_v, int32_t *perm, int32_t *kofc,
                  int32_t *labels, int32_t *trips, int32_t *label_hist, int32_t *swap_acc,
                  int32_t *swap_try, double *acc, int n_bins, void *e_series_v, int32_t *q_series,
                  uint64_t step0, int64_t n_steps, int multispin)
{
    const int8_t *J = jtype ? NULL : (const int8_t *)Jv;
    const double *Jf = jtype ? (const double *)Jv : NULL;
    int32_t *ei = jtype ? NULL : (int32_t *)energies_v;
    double *ef = jtype ? (double *)energies_v : NULL;
    int32_t dims[3] = {L, L, L};
    orc_cb g;
    if (orc_cb_init(&g, ndim, dims)) return -2;
    if (code != 0 && code != 2) return -3;
    if (multispin && (jtype || m > 32)) return -4;
    const int z = 2 * ndim;
    const int64_t N = g.N;
    uint64_t *thr = (uint64_t *)malloc(sizeof(uint64_t) * m * 16);
    for (int k = 0; k < m; ++k)
        for (int v = -z; v <= z; ++v) {
            double p = code == 0 ? (v <= 0 ? 1.0 : exp(-betas[k] * 2.0 * v))
                                 : 1.0 / (1.0 + exp(-2.0 * betas[k] * v));
            thr[k * 16 + v + z] = orc_thr_p(p);
        }
    int64_t nb[6];
    int64_t next_edge = 1;
    int ebin = 0;
    /* working energies: exact for integers, the running FP64 sums otherwise */
    double *E = (double *)malloc(sizeof(double) * 2 * m);
    for (int c = 0; c < 2 * m; ++c) E[c] = jtype ? ef[c] : (double)ei[c];
    for (int64_t t = 0; t < n_steps; ++t) {
        const uint64_t step = step0 + (uint64_t)t;
        if (t >= next_edge) {
            ++ebin;
            next_edge = ((int64_t)1 << (ebin + 1)) - 1;
            if (jtype)
                for (int c = 0; c < 2 * m; ++c) E[c] = orc_ea_energy_f64(spins + (int64_t)c * N, Jf, ndim, L);
        }
        if (multispin) {
            /* multi-spin mapping: the set's planes are shared by its configurations */
            uint32_t planes[32];
            for (int r = 0; r < 2; ++r) {
                uint64_t sw = sweeps[r * m];
                for (int sp = 0; sp < sweeps_per_step; ++sp, ++sw)
                    for (int col = 0; col < 2; ++col)
                        for (int64_t kk = 0; kk < g.half; ++kk) {
                            for (int q = 0; q < 8; ++q)
                                orc_draw_block(k0, k1, (uint32_t)r, sw, col, ORC_TAG_EA,
                                               (uint32_t)(8 * kk + q), planes + 4 * q);
                            int64_t i = orc_cb_site(&g, kk, col, nb);
                            for (int c = 0; c < m; ++c) {
                                int8_t *S = spins + ((int64_t)r * m + c) * N;
                                const uint64_t *T = thr + kofc[r * m + c] * 16;
                                uint32_t x = 0;
                                for (int t = 0; t < 32; ++t) x |= ((planes[t] >> c) & 1u) << (31 - t);
                                int si = S[i];
                                int h = 0;
                                for (int ax = 0; ax < ndim; ++ax)
                                    h += J[ax * N + i] * S[nb[2 * ax]] +
                                         J[ax * N + nb[2 * ax + 1]] * S[nb[2 * ax + 1]];
                                if (code == 0) {
                                    int aa = si * h;
                                    if (aa <= 0 || (uint64_t)x < T[aa + z]) {
                                        S[i] = (int8_t)(-si);
                                        E[r * m + c] += 2 * aa;
                                    }
                                } else {
                                    int nw = (uint64_t)x < T[h + z] ? 1 : -1;
                                    if (nw != si) {
                                        E[r * m + c] += 2 * si * h;
                                        S[i] = (int8_t)nw;
                                    }
                                }
                            }
                        }
                for (int c = 0; c < m; ++c) sweeps[r * m + c] = sw;
            }
        }
        for (int r = 0; r < 2 && !multispin; ++r)
            for (int c = 0; c < m; ++c) {
                int8_t *S = spins + ((int64_t)r * m + c) * N;
                const uint64_t *T = thr + kofc[r * m + c] * 16;
                uint64_t *sw = &sweeps[r * m + c];
                for (int sp = 0; sp < sweeps_per_step; ++sp) {
                    for (int col = 0; col < 2; ++col) {
                        uint32_t blk[4];
                        for (int64_t kk = 0; kk < g.half; ++kk) {
                            if ((kk & 3) == 0)
                                orc_draw_block(k0, k1, (uint32_t)(r * m + c), *sw, col, ORC_TAG_EA,
                                               (uint32_t)(kk >> 2), blk);
                            int64_t i = orc_cb_site(&g, kk, col, nb);
                            int si = S[i];
                            uint32_t x = blk[kk & 3];
                            if (jtype) {
                                const double beta = betas[kofc[r * m + c]];
                                double h = 0.0;
                                for (int ax = 0; ax < ndim; ++ax) {
                                    h += Jf[ax * N + i] * (double)S[nb[2 * ax]];
                                    h += Jf[ax * N + nb[2 * ax + 1]] * (double)S[nb[2 * ax + 1]];
                                }
                                double u = (double)x * 0x1p-32;
                                if (code == 0) {
                                    double dE = 2.0 * (double)si * h;
                                    if (dE <= 0.0 || u < orc_det_exp(-beta * dE)) {
                                        S[i] = (int8_t)(-si);
                                        E[r * m + c] += dE;
                                    }
                                } else {
                                    int nw = u < 1.0 / (1.0 + orc_det_exp(-2.0 * beta * h)) ? 1 : -1;
                                    if (nw != si) {
                                        E[r * m + c] += 2.0 * (double)si * h;
                                        S[i] = (int8_t)nw;
                                    }
                                }
                                continue;
                            }
                            int h = 0;
                            for (int ax = 0; ax < ndim; ++ax)
                                h += J[ax * N + i] * S[nb[2 * ax]] +
                                     J[ax * N + nb[2 * ax + 1]] * S[nb[2 * ax + 1]];
                            if (code == 0) {
                                int aa = si * h;
                                if (aa <= 0 || (uint64_t)x < T[aa + z]) {
                                    S[i] = (int8_t)(-si);
                                    E[r * m + c] += 2 * aa;
                                }
                            } else {
                                int nw = (uint64_t)x < T[h + z] ? 1 : -1;
                                if (nw != si) {
                                    E[r * m + c] += 2 * si * h;
                                    S[i] = (int8_t)nw;
                                }
                            }
                        }
                    }
                    *sw += 1;
                }
            }
        for (int r = 0; r < 2; ++r) {
            int32_t *p = perm + r * m, *ko = kofc + r * m, *lab = labels + r * m;
            const double *Er = E + r * m;
            uint32_t blk[4] = {0, 0, 0, 0};
            for (int k = 0; k < m - 1; ++k) {
                if ((k & 3) == 0)
                    orc_draw_block(k0, k1, (uint32_t)r, step, 0, ORC_TAG_PT, (uint32_t)(k >> 2), blk);
                int c1 = p[k], c2 = p[k + 1];
                double x = (betas[k] - betas[k + 1]) * (Er[c1] - Er[c2]);
                swap_try[r * (m - 1) + k] += 1;
                if (x >= 0.0 || (double)blk[k & 3] * 0x1p-32 < orc_det_exp(x)) {
                    p[k] = c2;
                    p[k + 1] = c1;
                    ko[c2] = k;
                    ko[c1] = k + 1;
                    swap_acc[r * (m - 1) + k] += 1;
                }
            }
            int hot = p[0], cold = p[m - 1];
            if (lab[hot] == 2) trips[r] += 1;
            lab[hot] = 1;
            if (lab[cold] == 1) trips[r] += 1;
            lab[cold] = 2;
            for (int k = 0; k < m; ++k) {
                int l = lab[p[k]];
                if (l == 1) label_hist[(r * m + k) * 2] += 1;
                else if (l == 2) label_hist[(r * m + k) * 2 + 1] += 1;
            }
        }
        int bin = 0;
        while (((uint64_t)t + 1) >> (bin + 1)) ++bin;
        for (int k = 0; k < m; ++k) {
            int ca = perm[k], cb = perm[m + k];
            const int8_t *sa = spins + (int64_t)ca * N, *sb = spins + ((int64_t)m + cb) * N;
            int64_t q = 0, ql = 0;
            double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
            for (int64_t i = 0; i < N; ++i) {
                int qi = sa[i] * sb[i];
                q += qi;
                int64_t rest = i;
                for (int ax = 0; ax < ndim; ++ax) {
                    int x = (int)(rest % L);
                    rest /= L;
                    int64_t j = ea_nbr_fwd(i, ax, L);
                    ql += qi * sa[j] * sb[j];
                    re[ax] += qi * cos(2.0 * M_PI * x / L);
                    im[ax] += qi * sin(2.0 * M_PI * x / L);
                }
            }
            double qk2 = 0.0;
            for (int ax = 0; ax < ndim; ++ax)
                qk2 += (re[ax] * re[ax] + im[ax] * im[ax]) / ((double)N * (double)N);
            qk2 /= ndim;
            double Ea = E[ca], Eb = E[m + cb];
            if (bin < n_bins) {
                double *o = acc + ((int64_t)bin * m + k) * 7;
                double qq = (double)q / (double)N;
                o[0] += 0.5 * (Ea + Eb) / (double)N;
                o[1] += (double)ql / (double)(ndim * N);
                o[2] += fabs(qq);
                o[3] += qq * qq;
                o[4] += qq * qq * qq * qq;
                o[5] += qk2;
                o[6] += 1.0;
            }
            if (e_series_v && jtype) {
                ((double *)e_series_v)[((int64_t)0 * m + k) * n_steps + t] = Ea;
                ((double *)e_series_v)[((int64_t)1 * m + k) * n_steps + t] = Eb;
            } else if (e_series_v) {
                ((int32_t *)e_series_v)[((int64_t)0 * m + k) * n_steps + t] = (int32_t)Ea;
                ((int32_t *)e_series_v)[((int64_t)1 * m + k) * n_steps + t] = (int32_t)Eb;
            }
            if (q_series) q_series[(int64_t)k * n_steps + t] = (int32_t)q;
        }
    }
    for (int c = 0; c < 2 * m; ++c) {
        if (jtype) ef[c] = E[c];
        else ei[c] = (int32_t)E[c];
    }
    free(E);
    free(thr);
    return 0;
}

/* Random start of every configuration: site i of configuration (r, c) takes
 * bit 31 of word i&3 of block i>>2 at (sweep 0, colour 0, TAG_INIT, stream r*m + c). */
void orc_ea_init(int8_t *spins, int m, int64_t N, uint32_t k0, uint32_t k1)
{
    uint32_t w[4];
    for (int rc = 0; rc < 2 * m; ++rc)
        for (int64_t i = 0; i < N; ++i) {
            if ((i & 3) == 0) orc_draw_block(k0, k1, (uint32_t)rc, 0, 0, ORC_TAG_INIT, (uint32_t)(i >> 2), w);
            spins[(int64_t)rc * N + i] = (w[i & 3] >> 31) ? 1 : -1;
        }
}
