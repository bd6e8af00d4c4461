/*
 * ORACLE (test infrastructure only) -- CPU restatement of the reference's
 * schedule interpreter for Ising (q = 0: spins +-1) and q-state Potts
 * (q >= 2: states 0..q-1) chains on any lattice given by its neighbour table
 * nbr[N][z] in pair form (column 2a = +e_a, 2a+1 = -e_a; critmc/lattice.py:
 * 129-134, 173-213), the bit-exact target of the sequential GPU kernels.
 *
 * Restated from critmc/kernels/ising.py:47-107, 110-170, 173-237 and
 * critmc/kernels/potts.py:44-167, 170-235, in the reference's own site orders:
 *   metropolis (0)         sites 0..N-1 in turn (typewriter)
 *   metropolis-random (1)  N attempts at uniformly drawn sites
 *   heatbath (2)           sites 0..N-1 in turn
 *   wolff (3), sw (4)      as replay_cluster.c, neighbours from the table:
 *                          bond b = a N + i joins i and nbr[i][2a], a < z/2
 * Each repetition of any row takes one value of the chain's sweep counter.
 * Uniforms (tag ORC_TAG_SEQ, key and stream of the chain): attempt t of a
 * local sweep uses block t of colour field = update code; its words are
 *   Ising metropolis:    w0 = acceptance x (accept iff x < ceil(p 2^32))
 *   Ising random:        w0 -> site floor(w0 N / 2^32), w1 = acceptance
 *   Ising heat bath:     w0: new spin +1 iff w0 < ceil(p_up 2^32)
 *   Potts metropolis:    w0 -> proposal floor(w0 (q-1) / 2^32), +1 if >= old;
 *                        w1 = acceptance (delta <= 0, or w1 < ceil(e^{-beta delta} 2^32))
 *   Potts random:        w0 -> site, w1 -> proposal, w2 = acceptance
 *   Potts heat bath:     u = w0 2^-32; new = first k with u sum_j w_j < sum_{j<=k} w_j,
 *                        w_k = exp(beta n_k) (FP64, in the reference's order), else q-1
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "philox.h"

static uint64_t orc_tab_thr(double p)
{
    if (!(p < 1.0)) return 1ull << 32;
    if (!(p > 0.0)) return 0;
    return (uint64_t)ceil(p * 4294967296.0);
}

static void orc_seq_block(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep, int field,
                          uint32_t t, uint32_t w[4])
{
    orc_draw_block(k0, k1, stream, sweep, field, ORC_TAG_SEQ, t, w);
}

static uint32_t orc_tab_cl_word(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep,
                                int field, int64_t index)
{
    uint32_t w[4];
    orc_draw_block(k0, k1, stream, sweep, field, ORC_TAG_CLUSTER, (uint32_t)(index >> 2), w);
    return w[index & 3];
}

/* Bond energy (Ising: -sum s_i s_j; Potts: -#equal) over the forward bonds. */
static int64_t orc_tab_energy(const int8_t *s, int64_t N, int z, const int32_t *nbr, int q)
{
    int64_t e = 0;
    for (int64_t i = 0; i < N; ++i)
        for (int a = 0; a < z / 2; ++a) {
            const int32_t j = nbr[i * z + 2 * a];
            e -= q == 0 ? (int64_t)s[i] * s[j] : (int64_t)(s[i] == s[j]);
        }
    return e;
}

typedef struct {
    int q, z;
    int64_t N;
    const int32_t *nbr;
    uint32_t k0, k1, stream;
    uint64_t met[2 * 25], hb[25]; /* Ising: index 2 (a + z) + spin bit, a in -z..z; Potts: delta */
    double wts[13];               /* Potts heat bath: exp(beta n), n = 0..z */
    uint64_t t_add;
    int64_t *nk;
} orc_tab;

static int64_t orc_tab_local(orc_tab *c, int8_t *s, int code, uint64_t sweep, int64_t *e,
                             int64_t *m, int64_t *counts)
{
    const int64_t N = c->N;
    const int z = c->z, q = c->q;
    int64_t accepted = 0;
    for (int64_t t = 0; t < N; ++t) {
        uint32_t w[4];
        orc_seq_block(c->k0, c->k1, c->stream, sweep, code, (uint32_t)t, w);
        int64_t i = t;
        int o = 0; /* word offset after the site draw */
        if (code == 1) {
            i = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
            o = 1;
        }
        const int32_t *nb = c->nbr + i * z;
        if (q == 0) {
            int field = 0;
            for (int k = 0; k < z; ++k) field += s[nb[k]];
            const int si = s[i];
            if (code == 2) {
                const int nw = (uint64_t)w[0] < c->hb[field + z] ? 1 : -1;
                if (nw != si) {
                    *e += 2 * si * field;
                    *m += 2 * nw;
                    s[i] = (int8_t)nw;
                    ++accepted;
                }
            } else {
                const int a = si * field;
                if ((uint64_t)w[o] < c->met[2 * (a + z) + ((si + 1) >> 1)]) {
                    s[i] = (int8_t)(-si);
                    *e += 2 * a;
                    *m -= 2 * si;
                    ++accepted;
                }
            }
        } else {
            const int old = s[i];
            if (code == 2) {
                for (int k = 0; k < q; ++k) c->nk[k] = 0;
                for (int k = 0; k < z; ++k) c->nk[s[nb[k]]] += 1;
                double total = 0.0;
                for (int k = 0; k < q; ++k) total += c->wts[c->nk[k]];
                const double r = ((double)w[0] * 0x1p-32) * total;
                int nw = q - 1;
                double cum = 0.0;
                for (int k = 0; k < q; ++k) {
                    cum += c->wts[c->nk[k]];
                    if (r < cum) {
                        nw = k;
                        break;
                    }
                }
                if (nw != old) {
                    *e += c->nk[old] - c->nk[nw];
                    counts[old] -= 1;
                    counts[nw] += 1;
                    s[i] = (int8_t)nw;
                    ++accepted;
                }
            } else {
                int nw = (int)(((uint64_t)w[o] * (uint32_t)(q - 1)) >> 32);
                if (nw >= old) ++nw;
                int n_old = 0, n_new = 0;
                for (int k = 0; k < z; ++k) {
                    const int v = s[nb[k]];
                    n_old += v == old;
                    n_new += v == nw;
                }
                const int delta = n_old - n_new;
                if (delta <= 0 || (uint64_t)w[o + 1] < c->met[delta]) {
                    s[i] = (int8_t)nw;
                    counts[old] -= 1;
                    counts[nw] += 1;
                    *e += delta;
                    ++accepted;
                }
            }
        }
    }
    return accepted;
}

static int64_t orc_tab_wolff(orc_tab *c, int8_t *s, uint64_t sweep, int32_t *queue)
{
    const int64_t N = c->N;
    const int z = c->z, q = c->q;
    const int64_t seed =
        (int64_t)(((uint64_t)orc_tab_cl_word(c->k0, c->k1, c->stream, sweep, 2, 0) * (uint64_t)N) >> 32);
    const int a = s[seed];
    int b;
    if (q == 0) {
        b = -a;
    } else {
        b = (int)(((uint64_t)orc_tab_cl_word(c->k0, c->k1, c->stream, sweep, 2, 1) * (uint32_t)(q - 1)) >> 32);
        if (b >= a) ++b;
    }
    s[seed] = (int8_t)b;
    queue[0] = (int32_t)seed;
    int64_t head = 0, tail = 1;
    while (head < tail) {
        const int64_t k = queue[head++];
        for (int col = 0; col < z; ++col) {
            const int64_t j = c->nbr[k * z + col];
            if (s[j] != a) continue;
            const int64_t bond = (int64_t)(col >> 1) * N + ((col & 1) ? j : k);
            if ((uint64_t)orc_tab_cl_word(c->k0, c->k1, c->stream, sweep, 0, bond) >= c->t_add) continue;
            s[j] = (int8_t)b;
            queue[tail++] = (int32_t)j;
        }
    }
    return tail;
}

static int32_t orc_tab_find(int32_t *parent, int32_t i)
{
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
}

static void orc_tab_sw(orc_tab *c, int8_t *s, uint64_t sweep, int32_t *parent, int32_t *size,
                       int64_t *c2, int64_t *ncl)
{
    const int64_t N = c->N;
    for (int64_t i = 0; i < N; ++i) parent[i] = (int32_t)i;
    for (int a = 0; a < c->z / 2; ++a)
        for (int64_t i = 0; i < N; ++i) {
            const int64_t j = c->nbr[i * c->z + 2 * a];
            if (s[i] != s[j]) continue;
            if ((uint64_t)orc_tab_cl_word(c->k0, c->k1, c->stream, sweep, 0, (int64_t)a * N + i) >= c->t_add)
                continue;
            const int32_t ri = orc_tab_find(parent, (int32_t)i), rj = orc_tab_find(parent, (int32_t)j);
            if (ri == rj) continue;
            if (ri < rj) parent[rj] = ri;
            else parent[ri] = rj;
        }
    for (int64_t i = 0; i < N; ++i) size[i] = 0;
    for (int64_t i = 0; i < N; ++i) size[orc_tab_find(parent, (int32_t)i)] += 1;
    *c2 = 0;
    *ncl = 0;
    for (int64_t i = 0; i < N; ++i) {
        const int32_t r = orc_tab_find(parent, (int32_t)i);
        if (r == i) {
            *c2 += (int64_t)size[i] * size[i];
            *ncl += 1;
        }
        const uint32_t x = orc_tab_cl_word(c->k0, c->k1, c->stream, sweep, 1, r);
        if (c->q == 0) {
            if ((x >> 31) == 0) s[i] = (int8_t)(-s[i]);
        } else {
            s[i] = (int8_t)(((uint64_t)x * (uint32_t)c->q) >> 32);
        }
    }
}

/* Run `schedule` n_meas times.  Ising: e = E_b, mc[0] = M; Potts: e = E,
 * mc[0..q-1] = occupation counts (all running, in/out).  Records e_out[n_meas],
 * mc_out[n_meas][q or 1] and aux_out[n_meas][4] = (accepted, Wolff sizes,
 * sum |C|^2, clusters of the last Swendsen-Wang update); *sweep advances by one
 * per repetition of any row.  Returns 0, or < 0 for bad arguments. */
int orc_table_run(int8_t *s, int64_t N, int z, const int32_t *nbr, int q, double beta, double h,
                  uint32_t k0, uint32_t k1, uint32_t stream, uint64_t *sweep,
                  const int64_t *schedule, int n_rows, int64_t n_meas, int64_t *e, int64_t *mc,
                  int64_t *e_out, int64_t *mc_out, int64_t *aux_out)
{
    if (N < 1 || z < 2 || z > 12 || (z & 1) || (q != 0 && (q < 2 || q > 127))) return -1;
    orc_tab c;
    memset(&c, 0, sizeof c);
    c.q = q;
    c.z = z;
    c.N = N;
    c.nbr = nbr;
    c.k0 = k0;
    c.k1 = k1;
    c.stream = stream;
    if (q == 0) {
        for (int a = -z; a <= z; ++a) {
            for (int t = 0; t < 2; ++t) {
                const double de = 2.0 * a + 2.0 * h * (2 * t - 1);
                c.met[2 * (a + z) + t] = orc_tab_thr(de <= 0.0 ? 1.0 : exp(-beta * de));
            }
            c.hb[a + z] = orc_tab_thr(1.0 / (1.0 + exp(-2.0 * beta * (a + h))));
        }
        c.t_add = orc_tab_thr(1.0 - exp(-2.0 * beta));
    } else {
        for (int d = 0; d <= z; ++d) {
            c.met[d] = orc_tab_thr(exp(-beta * d));
            c.wts[d] = exp(beta * d);
        }
        c.t_add = orc_tab_thr(1.0 - exp(-beta));
    }
    const int nmc = q == 0 ? 1 : q;
    int64_t *nk = (int64_t *)malloc(sizeof(int64_t) * (q > 0 ? q : 1));
    int32_t *w1 = (int32_t *)malloc(sizeof(int32_t) * N), *w2 = (int32_t *)malloc(sizeof(int32_t) * N);
    c.nk = nk;
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t acc = 0, wsize = 0, c2 = 0, ncl = 0;
        for (int r = 0; r < n_rows; ++r) {
            const int code = (int)schedule[2 * r];
            for (int64_t rep = 0; rep < schedule[2 * r + 1]; ++rep) {
                const uint64_t sw = (*sweep)++;
                if (code <= 2) {
                    acc += orc_tab_local(&c, s, code, sw, e, q == 0 ? mc : NULL, mc);
                } else if (code == 3) {
                    wsize += orc_tab_wolff(&c, s, sw, w1);
                    *e = orc_tab_energy(s, N, z, nbr, q);
                } else {
                    orc_tab_sw(&c, s, sw, w1, w2, &c2, &ncl);
                    *e = orc_tab_energy(s, N, z, nbr, q);
                }
                if (code >= 3) { /* recount the order parameter after a cluster move */
                    for (int k = 0; k < nmc; ++k) mc[k] = 0;
                    for (int64_t i = 0; i < N; ++i) {
                        if (q == 0) mc[0] += s[i];
                        else mc[s[i]] += 1;
                    }
                }
            }
        }
        e_out[meas] = *e;
        for (int k = 0; k < nmc; ++k) mc_out[meas * nmc + k] = mc[k];
        aux_out[4 * meas + 0] = acc;
        aux_out[4 * meas + 1] = wsize;
        aux_out[4 * meas + 2] = c2;
        aux_out[4 * meas + 3] = ncl;
    }
    free(nk);
    free(w1);
    free(w2);
    return 0;
}
