/*
 * ORACLE (test infrastructure only) -- CPU restatement of the Wolff and
 * Swendsen-Wang updates of the Ising (q = 0: spins +-1) and q-state Potts
 * (q >= 2: states 0..q-1) models, bit-exact target of the GPU kernels.
 *
 * Restated from the reference: critmc/kernels/ising.py:110-170 and
 * critmc/kernels/potts.py:106-167.  One sweep-counter value per update;
 * uniforms at tag ORC_TAG_CLUSTER, key and stream of the chain:
 *   bond b = a N + i (i and its +e_a neighbour) is active iff the two states
 *   agree and word b&3 of block b>>2 (colour field 0) is below t_add
 *   (= ceil(p 2^32), p = 1 - exp(-2 beta) for Ising, 1 - exp(-beta) for Potts);
 *   Swendsen-Wang: clusters are the components of the active bonds, named by
 *   their smallest site r; with x = word r&3 of block r>>2 (colour field 1)
 *   an Ising cluster flips iff x's top bit is clear (u < 1/2) and a Potts
 *   cluster takes state floor(x q / 2^32);
 *   Wolff: seed = floor(x0 N / 2^32), x0 = word 0 of block 0 (colour field 2);
 *   the new state is -a (Ising) or randbelow(q - 1) from word 1, +1 if >= a
 *   (Potts); the cluster is the seed's component among sites in the seed's
 *   state a, recoloured as members join (breadth-first).
 */
#include <stdint.h>
#include <stdlib.h>

#include "philox.h"

typedef struct {
    int d;
    int64_t L[3], stride[3], N;
} orc_clg;

static void orc_clg_init(orc_clg *g, int ndim, const int32_t *dims)
{
    g->d = ndim;
    g->N = 1;
    for (int a = 0; a < 3; ++a) {
        g->L[a] = a < ndim ? dims[a] : 1;
        g->stride[a] = g->N;
        g->N *= g->L[a];
    }
}

static int64_t orc_clg_nbr(const orc_clg *g, int64_t i, int a, int fwd)
{
    int64_t x = (i / g->stride[a]) % g->L[a];
    if (fwd) return x + 1 == g->L[a] ? i - (g->L[a] - 1) * g->stride[a] : i + g->stride[a];
    return x == 0 ? i + (g->L[a] - 1) * g->stride[a] : i - g->stride[a];
}

static uint32_t orc_cl_word(uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep, int field,
                            int64_t index)
{
    uint32_t w[4];
    orc_draw_block(k0, k1, stream, sweep, field, ORC_TAG_CLUSTER, (uint32_t)(index >> 2), w);
    return w[index & 3];
}

static int32_t orc_cl_find(int32_t *parent, int32_t i)
{
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
}

int64_t orc_cluster_wolff(int8_t *s, int ndim, const int32_t *dims, int q, uint64_t t_add,
                          uint32_t k0, uint32_t k1, uint32_t stream, uint64_t sweep,
                          int32_t *queue)
{
    orc_clg g;
    orc_clg_init(&g, ndim, dims);
    const int64_t seed = (int64_t)(((uint64_t)orc_cl_word(k0, k1, stream, sweep, 2, 0) * (uint64_t)g.N) >> 32);
    const int a = s[seed];
    int b;
    if (q == 0) {
        b = -a;
    } else {
        b = (int)(((uint64_t)orc_cl_word(k0, k1, stream, sweep, 2, 1) * (uint32_t)(q - 1)) >> 32);
        if (b >= a) ++b;
    }
    s[seed] = (int8_t)b;
    queue[0] = (int32_t)seed;
    int64_t head = 0, tail = 1;
    while (head < tail) {
        const int64_t k = queue[head++];
        for (int c = 0; c < 2 * g.d; ++c) {
            const int ax = c >> 1;
            const int64_t j = orc_clg_nbr(&g, k, ax, (c & 1) == 0);
            if (s[j] != a) continue;
            const int64_t bond = (int64_t)ax * g.N + ((c & 1) ? j : k);
            if ((uint64_t)orc_cl_word(k0, k1, stream, sweep, 0, bond) >= t_add) continue;
            s[j] = (int8_t)b;
            queue[tail++] = (int32_t)j;
        }
    }
    return tail;
}

void orc_cluster_sw(int8_t *s, int ndim, const int32_t *dims, int q, uint64_t t_add, uint32_t k0,
                    uint32_t k1, uint32_t stream, uint64_t sweep, int32_t *parent, int32_t *size,
                    int64_t *c2, int64_t *ncl)
{
    orc_clg g;
    orc_clg_init(&g, ndim, dims);
    for (int64_t i = 0; i < g.N; ++i) parent[i] = (int32_t)i;
    for (int a = 0; a < g.d; ++a)
        for (int64_t i = 0; i < g.N; ++i) {
            const int64_t j = orc_clg_nbr(&g, i, a, 1);
            if (s[i] != s[j]) continue;
            if ((uint64_t)orc_cl_word(k0, k1, stream, sweep, 0, (int64_t)a * g.N + i) >= t_add) continue;
            int32_t ri = orc_cl_find(parent, (int32_t)i), rj = orc_cl_find(parent, (int32_t)j);
            if (ri == rj) continue;
            if (ri < rj) parent[rj] = ri; /* the smaller site is the root */
            else parent[ri] = rj;
        }
    for (int64_t i = 0; i < g.N; ++i) size[i] = 0;
    for (int64_t i = 0; i < g.N; ++i) size[orc_cl_find(parent, (int32_t)i)] += 1;
    *c2 = 0;
    *ncl = 0;
    for (int64_t i = 0; i < g.N; ++i) {
        const int32_t r = orc_cl_find(parent, (int32_t)i);
        if (r == i) {
            *c2 += (int64_t)size[i] * size[i];
            *ncl += 1;
        }
        const uint32_t x = orc_cl_word(k0, k1, stream, sweep, 1, r);
        if (q == 0) {
            if ((x >> 31) == 0) s[i] = (int8_t)(-s[i]);
        } else {
            s[i] = (int8_t)(((uint64_t)x * (uint32_t)q) >> 32);
        }
    }
}
