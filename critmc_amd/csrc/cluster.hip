// Cluster updates of the Ising and q-state Potts models: Swendsen-Wang and
// Wolff (critmc/kernels/ising.py:110-170, critmc/kernels/potts.py:106-167),
// batched over replicas of one lattice.  Ising spins are int8 +-1 (q = 0
// below), Potts states int8 0..q-1.
//
// Random numbers (tag TAG_CLUSTER, key and stream of the chain, one sweep
// counter value per cluster update):
//   bond b = a N + i joins site i and its +e_a neighbour; it is active iff the
//   two states agree and word b&3 of block b>>2 (colour field 0) is below
//   ceil(p 2^32), p = 1 - exp(-2 beta) (Ising) or 1 - exp(-beta) (Potts), the
//   reference's p_add;
//   Swendsen-Wang: with x = word r&3 of block r>>2 (colour field 1) for the
//   cluster whose smallest site is r, an Ising cluster flips iff x's top bit
//   is clear (u < 1/2) and a Potts cluster takes state floor(x q / 2^32);
//   Wolff: the seed is floor(x N / 2^32), x = word 0 of block 0 (colour
//   field 2); the new state is -a (Ising) or randbelow(q - 1) from word 1,
//   +1 if >= a (Potts); the cluster is the seed's component among sites of
//   the seed's state a joined by active bonds.
// Every bond has its own uniform, so a cluster does not depend on the order in
// which it is explored: the device and a sequential host replay build the
// same clusters.
//
// Swendsen-Wang runs as four launches over all replicas: parent init, union of
// the active bonds (lock-free: a root is only ever hooked below a smaller
// root, with atomicCAS, so each component's root ends as its smallest site),
// flatten + cluster sizes, flip + (sum |C|^2, cluster count).  Wolff runs one
// workgroup per replica: a breadth-first search whose frontier lives in a
// per-replica queue, members flipped as they join (the reference's order of
// effects), `count` clusters per launch.
//
// A schedule mixing local and cluster rows runs row by row: local rows through
// cm_ising_run / cm_potts_run, cluster rows here, then the energy and
// magnetisation or state counts are recounted (cm_ising_measure /
// cm_potts_measure), as the reference recounts after Swendsen-Wang.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cb_geom.hpp"
#include "cluster_core.hpp"
#include "common.hpp"
#include "philox.hpp"
#include "seq_core.hpp"

extern "C" {
int cm_ising_run(void* spins, int32_t layout, int32_t ndim, const int32_t* dims, int32_t n_rep,
                 const double* betas, double h, const uint32_t* keys, const uint32_t* stream_ids,
                 uint64_t* sweeps, const uint64_t* sweeps_host, const int64_t* schedule,
                 int32_t n_rows, int64_t n_meas, int32_t* e_run, int32_t* m_run, int32_t* e_out,
                 int32_t* m_out, int32_t* acc_out, void* workspace, int64_t workspace_bytes,
                 int32_t form, int32_t block, int32_t grid, const int64_t* n_meas_rep,
                 const int32_t* prepared, void* stream);
int cm_ising_measure(const void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                     int32_t* e_run, int32_t* m_run, void* stream);
int cm_ising_prepare(int32_t ndim, const int32_t* dims, int32_t n_rep, const double* betas,
                     double h, const int64_t* schedule, int32_t n_rows, void* workspace,
                     int64_t workspace_bytes, int32_t* info, void* stream);
int64_t cm_ising_workspace_bytes(int32_t n_rep);
int cm_potts_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
                 const double* betas, const uint32_t* keys, const uint32_t* stream_ids,
                 uint64_t* sweeps, const int64_t* schedule, int32_t n_rows, int64_t n_meas,
                 int32_t* e_run, int32_t* counts_run, int32_t* e_out, int32_t* acc_out,
                 int32_t* counts_out, void* workspace, int64_t workspace_bytes, int32_t block,
                 int32_t grid, const int64_t* n_meas_rep, void* stream);
int cm_potts_measure(const void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                     int32_t q, int32_t* e_run, int32_t* counts_run, void* stream);
int64_t cm_potts_workspace_bytes(int32_t n_rep);
}

namespace cm {

// `rows` applied n_meas times to every chain, one lane per chain.  With
// e_out != null each measurement records E (and M or counts, and aux) as the
// schedule interpreter of the reference does; otherwise accepted moves are
// added to a.acc.  Each repetition of a row takes one sweep-counter value.
__global__ __launch_bounds__(64) void seq_run(SeqArgs a, SeqRows rows, int64_t n_meas,
                                              int32_t* e_out, int32_t* mc_out, int64_t* aux_out)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= a.R) return;
    int8_t* s = a.spins + (int64_t)rep * a.N;
    const uint64_t* tab = a.tab + (int64_t)rep * SEQ_TAB;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    int32_t* cnt = a.q ? a.counts + (int64_t)rep * a.q : nullptr;
    uint64_t sweep = a.sweeps[rep];
    int e = a.e_run[rep];
    int m = a.q ? 0 : a.m_run[rep];
    int64_t total = 0;
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        int64_t accepted = 0;
        for (int r = 0; r < rows.n; ++r)
            for (int k = 0; k < rows.count[r]; ++k, ++sweep)
                accepted += a.q ? seq_potts_sweep(a, s, tab, cnt, rows.code[r], sweep, k0, k1, sid, e)
                                : seq_ising_sweep(a, s, tab, rows.code[r], sweep, k0, k1, sid, e, m);
        total += accepted;
        if (e_out) {
            const int64_t o = (int64_t)rep * n_meas + meas;
            e_out[o] = e;
            if (a.q) {
                for (int c = 0; c < a.q; ++c) mc_out[o * a.q + c] = cnt[c];
            } else {
                mc_out[o] = m;
            }
            aux_out[4 * o + 0] = accepted;
            aux_out[4 * o + 1] = 0;
            aux_out[4 * o + 2] = 0;
            aux_out[4 * o + 3] = 0;
        }
    }
    a.sweeps[rep] = sweep;
    a.e_run[rep] = e;
    if (!a.q) a.m_run[rep] = m;
    if (a.acc) a.acc[rep] += total;
}

// Exact E (and M or counts) of every chain from the table: one workgroup per chain.
__global__ __launch_bounds__(256) void seq_measure(SeqArgs a)
{
    __shared__ long long red[2][4];
    const int rep = blockIdx.x;
    const int8_t* s = a.spins + (int64_t)rep * a.N;
    int32_t* cnt = a.q ? a.counts + (int64_t)rep * a.q : nullptr;
    if (a.q)
        for (int k = threadIdx.x; k < a.q; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    long long e = 0, m = 0;
    for (uint32_t i = threadIdx.x; i < a.N; i += blockDim.x) {
        const int si = s[i];
        const int32_t* nb = a.nbr + (int64_t)i * a.z;
        for (int ax = 0; ax < a.z / 2; ++ax) {
            const int sj = s[nb[2 * ax]];
            e -= a.q ? (long long)(si == sj) : (long long)(si * sj);
        }
        if (a.q) atomicAdd(cnt + si, 1);
        else m += si;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        m += __shfl_xor(m, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = e;
        red[1][threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long te = 0, tm = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            te += red[0][w];
            tm += red[1][w];
        }
        a.e_run[rep] = (int32_t)te;
        if (!a.q) a.m_run[rep] = (int32_t)tm;
    }
}

// potts.hip: the tables of cm_potts_run, and its launches without the upload
int potts_prepare(int32_t ndim, const int32_t* dims, int32_t n_rep, const double* betas,
                  void* workspace, int64_t workspace_bytes, void* stream);
int potts_rows(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
               const uint32_t* keys, const uint32_t* stream_ids, uint64_t* sweeps,
               const int64_t* schedule, int32_t n_rows, int64_t n_meas, int32_t* e_run,
               int32_t* counts_run, int32_t* e_out, int32_t* acc_out, int32_t* counts_out,
               void* workspace, int64_t workspace_bytes, int32_t block, int32_t grid,
               const int64_t* n_meas_rep, void* stream);
// ising.hip: whether cm_ising_run's automatic form only launches kernels
bool ising_rows_capturable(int32_t ndim, const int32_t* dims, int32_t n_rep);
}  // namespace cm

namespace cm {

struct ClArgs {
    int8_t* spins;          // [R][N] int8
    int32_t* parent;        // [R][N]
    int32_t* size;          // [R][N]
    int32_t* queue;         // [R][N]
    int32_t* mark;          // [R][N]
    int64_t* stats;         // [R][2]: sum |C|^2, clusters of the last Swendsen-Wang update
    int64_t* wsize;         // [R]: Wolff cluster sizes of the current measurement
    const uint64_t* t_add;  // [R]
    const uint32_t* keys;   // [R][2]
    const uint32_t* sids;   // [R]
    const uint64_t* sweeps; // [R]
    uint64_t sweep_add;
    ClGeo g;
    int R;
    int q;                  // 0: Ising +-1, else Potts states 0..q-1
};

// parent[i] = i, size[i] = 0, stats = 0
__global__ __launch_bounds__(256) void sw_init(ClArgs a)
{
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    int32_t* P = a.parent + (int64_t)rep * N;
    int32_t* Z = a.size + (int64_t)rep * N;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        P[i] = (int32_t)i;
        Z[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2) a.stats[2 * rep + threadIdx.x] = 0;
}

// One lane per Philox block of bonds (4 bonds).
__global__ __launch_bounds__(256) void sw_union(ClArgs a)
{
    const int rep = blockIdx.y;
    const ClGeo& g = a.g;
    const uint32_t N = g.N, nb = (uint32_t)cl_axes(g) * N;
    const int8_t* S = a.spins + (int64_t)rep * N;
    int32_t* P = a.parent + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
    for (uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x; 4 * blk < nb;
         blk += gridDim.x * blockDim.x) {
        const u32x4 r = philox4x32_10(u32x4{blk, (uint32_t)sweep, zc, sid}, k0, k1);
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t b = 4 * blk + t;
            if (b >= nb) break;
            if ((uint64_t)w[t] >= T) continue;
            int ax = 0;
            uint32_t i = b;
            while (i >= N) {  // at most three steps: z <= 8
                i -= N;
                ++ax;
            }
            const uint32_t j = cl_nbr_col(g, i, 2 * ax);
            if (S[i] != S[j]) continue;
            cl_unite(P, i, j);
        }
    }
}

// parent[i] = root(i); size[root] += 1.  Only lane i writes parent[i], and it
// writes a root, so the flip pass reads every site's root.  Near the critical
// point one cluster holds a large share of the sites, and one atomic add per
// site on its root serialises; so up to three times the wave takes the root of
// its first remaining lane and adds, in one atomic, the number of lanes that
// share it.  The sizes are integer sums, so they do not depend on the order.
__global__ __launch_bounds__(256) void sw_flatten(ClArgs a)
{
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    int32_t* P = a.parent + (int64_t)rep * N;
    int32_t* Z = a.size + (int64_t)rep * N;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < N; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;  // the wave stays together: i0 is uniform
        const bool site = i < N;
        uint32_t r = 0;
        if (site) {
            r = cl_find_ro(P, i);
            st_relaxed(P + i, r);
        }
        bool left = site;
#pragma unroll
        for (int pass = 0; pass < 3; ++pass) {
            const uint64_t rem = __ballot(left);
            if (rem == 0) break;
            const int first = __ffsll((long long)rem) - 1;
            const uint32_t r0 = (uint32_t)__shfl((int)r, first, 64);
            const bool mine = left && r == r0;
            const uint64_t grp = __ballot(mine);
            if (mine && (int)(threadIdx.x & 63) == first) atomicAdd(Z + r0, __popcll(grp));
            left = left && !mine;
        }
        if (left) atomicAdd(Z + r, 1);
    }
}

// Flip every cluster whose flip bit is set; sum |C|^2 and count the clusters.
__global__ __launch_bounds__(256) void sw_flip(ClArgs a)
{
    __shared__ long long red[2][4];
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    int8_t* S = a.spins + (int64_t)rep * N;
    const int32_t* P = a.parent + (int64_t)rep * N;
    const int32_t* Z = a.size + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    long long c2 = 0, ncl = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)P[i];
        if (r == i) {
            const long long z = Z[i];
            c2 += z * z;
            ncl += 1;
        }
        const uint32_t x = cl_word(k0, k1, sid, sweep, 1, r);
        if (a.q == 0) {
            if ((x >> 31) == 0) S[i] = (int8_t)(-S[i]);
        } else {
            S[i] = (int8_t)(((uint64_t)x * (uint32_t)a.q) >> 32);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c2 += __shfl_xor(c2, off, 64);
        ncl += __shfl_xor(ncl, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = c2;
        red[1][threadIdx.x >> 6] = ncl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t0 = 0, t1 = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            t0 += red[0][w];
            t1 += red[1][w];
        }
        if (t0) atomicAdd((unsigned long long*)&a.stats[2 * rep], (unsigned long long)t0);
        if (t1) atomicAdd((unsigned long long*)&a.stats[2 * rep + 1], (unsigned long long)t1);
    }
}

__global__ void cl_fill(int32_t* p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// `count` Wolff clusters per replica; one workgroup per replica.
__global__ __launch_bounds__(256) void wolff_run(ClArgs a, int count)
{
    __shared__ uint32_t s_head, s_tail;
    __shared__ int s_a, s_b;
    const int rep = blockIdx.x;
    const ClGeo& g = a.g;
    const uint32_t N = g.N;
    int8_t* S = a.spins + (int64_t)rep * N;
    int32_t* Q = a.queue + (int64_t)rep * N;
    int32_t* Mk = a.mark + (int64_t)rep * N;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    const int z = 2 * cl_axes(g);
    long long total = 0;
    for (int f = 0; f < count; ++f) {
        const uint64_t sweep = a.sweeps[rep] + a.sweep_add + (uint64_t)f;
        if (threadIdx.x == 0) {
            const u32x4 r = philox4x32_10(
                u32x4{0u, (uint32_t)sweep, ctr_z(sweep, 2, TAG_CLUSTER), sid}, k0, k1);
            const uint32_t seed = (uint32_t)(((uint64_t)r.x * N) >> 32);
            const int av = S[seed];
            int bv;
            if (a.q == 0) {
                bv = -av;
            } else {
                bv = (int)(((uint64_t)r.y * (uint32_t)(a.q - 1)) >> 32);
                if (bv >= av) ++bv;
            }
            S[seed] = (int8_t)bv;
            Mk[seed] = f;
            Q[0] = (int32_t)seed;
            s_a = av;
            s_b = bv;
            s_head = 0;
            s_tail = 1;
        }
        __syncthreads();
        const int s0 = s_a, s1 = s_b;
        const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
        while (true) {
            const uint32_t lo = s_head, hi = s_tail;
            // every lane has this level's bounds before any lane appends to the queue
            __syncthreads();
            if (lo == hi) break;
            // one (site, neighbour) pair per lane: a level's Philox draws
            // run side by side rather than z after each other in one lane
            const uint32_t npair = (hi - lo) * (uint32_t)z;
            for (uint32_t t = threadIdx.x; t < npair; t += blockDim.x) {
                const uint32_t k = (uint32_t)Q[lo + t / (uint32_t)z];
                const int c = (int)(t % (uint32_t)z);
                {
                    const int ax = c >> 1;
                    uint32_t x[3];
                    if (!g.nbr) cl_coords(g, k, x);
                    const uint32_t j = g.nbr ? (uint32_t)g.nbr[(int64_t)k * g.zt + c]
                                             : cl_nbr(g, k, x, ax, (c & 1) == 0);
                    if (S[j] != s0) continue;
                    const uint32_t b = (uint32_t)ax * N + ((c & 1) ? j : k);
                    const u32x4 r = philox4x32_10(u32x4{b >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
                    const uint32_t w = b & 3;
                    const uint32_t u = w == 0 ? r.x : (w == 1 ? r.y : (w == 2 ? r.z : r.w));
                    if ((uint64_t)u >= T) continue;
                    if (atomicExch(Mk + j, f) == f) continue;  // another lane added it
                    S[j] = (int8_t)s1;
                    Q[atomicAdd(&s_tail, 1u)] = (int32_t)j;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) s_head = hi;
            __syncthreads();
        }
        total += s_tail;
        __syncthreads();
    }
    if (threadIdx.x == 0) a.wsize[rep] += total;
}

// wolff_run with the replica's spins and a membership bitmask in LDS (Potts
// lattices of at most CL_WOLFF_LDS_SITES sites; Ising ones take
// wolff_ising_bits): the spin tests, claims and flips of the search are LDS
// operations, and only the queue stays in HBM.  The marks
// are cleared per cluster, so a claim is "bit j newly set".  The search, and
// with it every cluster, is that of wolff_run.
constexpr int64_t CL_WOLFF_LDS_SITES = 65536;

__global__ __launch_bounds__(256) void wolff_run_lds(ClArgs a, int count)
{
    extern __shared__ uint32_t wl[];
    __shared__ uint32_t s_head, s_tail;
    __shared__ int s_a, s_b;
    const int rep = blockIdx.x;
    const ClGeo& g = a.g;
    const uint32_t N = g.N;
    int8_t* Sg = a.spins + (int64_t)rep * N;
    int8_t* S = (int8_t*)wl;
    const uint32_t nmw = (N + 31) / 32;
    uint32_t* Mb = wl + (N + 3) / 4;
    int32_t* Q = a.queue + (int64_t)rep * N;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    const int z = 2 * cl_axes(g);
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) S[i] = Sg[i];
    long long total = 0;
    for (int f = 0; f < count; ++f) {
        for (uint32_t w = threadIdx.x; w < nmw; w += blockDim.x) Mb[w] = 0u;
        __syncthreads();
        const uint64_t sweep = a.sweeps[rep] + a.sweep_add + (uint64_t)f;
        if (threadIdx.x == 0) {
            const u32x4 r = philox4x32_10(
                u32x4{0u, (uint32_t)sweep, ctr_z(sweep, 2, TAG_CLUSTER), sid}, k0, k1);
            const uint32_t seed = (uint32_t)(((uint64_t)r.x * N) >> 32);
            const int av = S[seed];
            int bv;
            if (a.q == 0) {
                bv = -av;
            } else {
                bv = (int)(((uint64_t)r.y * (uint32_t)(a.q - 1)) >> 32);
                if (bv >= av) ++bv;
            }
            S[seed] = (int8_t)bv;
            Mb[seed >> 5] |= 1u << (seed & 31u);
            Q[0] = (int32_t)seed;
            s_a = av;
            s_b = bv;
            s_head = 0;
            s_tail = 1;
        }
        __syncthreads();
        const int s0 = s_a, s1 = s_b;
        const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
        while (true) {
            const uint32_t lo = s_head, hi = s_tail;
            // every lane has this level's bounds before any lane appends to the queue
            __syncthreads();
            if (lo == hi) break;
            // one (site, neighbour) pair per lane: a level's Philox draws
            // run side by side rather than z after each other in one lane
            const uint32_t npair = (hi - lo) * (uint32_t)z;
            for (uint32_t t = threadIdx.x; t < npair; t += blockDim.x) {
                const uint32_t k = (uint32_t)Q[lo + t / (uint32_t)z];
                const int c = (int)(t % (uint32_t)z);
                {
                    const int ax = c >> 1;
                    uint32_t x[3];
                    if (!g.nbr) cl_coords(g, k, x);
                    const uint32_t j = g.nbr ? (uint32_t)g.nbr[(int64_t)k * g.zt + c]
                                             : cl_nbr(g, k, x, ax, (c & 1) == 0);
                    if (S[j] != s0) continue;
                    const uint32_t b = (uint32_t)ax * N + ((c & 1) ? j : k);
                    const u32x4 r = philox4x32_10(u32x4{b >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
                    const uint32_t w = b & 3;
                    const uint32_t u = w == 0 ? r.x : (w == 1 ? r.y : (w == 2 ? r.z : r.w));
                    if ((uint64_t)u >= T) continue;
                    const uint32_t bit = 1u << (j & 31u);
                    if (atomicOr(&Mb[j >> 5], bit) & bit) continue;  // another lane added it
                    S[j] = (int8_t)s1;
                    Q[atomicAdd(&s_tail, 1u)] = (int32_t)j;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) s_head = hi;
            __syncthreads();
        }
        total += s_tail;
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) Sg[i] = S[i];
    if (threadIdx.x == 0) a.wsize[rep] += total;
}

// wolff_run for Ising lattices of at most CL_WOLFF_LDS_SITES sites with
// everything in LDS: the spins as bits (bit i of word i / 32 set for +1) and
// the queue as 16-bit site indices, 8 KB + 128 KB at 256^2.  A claim moves a
// site's bit away from the cluster's spin with one atomic AND or OR, whose old
// word says whether this lane made the change, so no membership mask is kept.
// Each level appends through its own counter of three that rotate, so one
// barrier per level separates the level's appends from the next level's
// reads: the counter for level L + 1 is cleared during level L, and it was
// last read at the start of level L - 1.  The search, and with it every
// cluster, is that of wolff_run.
__global__ __launch_bounds__(256) void wolff_ising_bits(ClArgs a, int count)
{
    extern __shared__ uint32_t wl[];
    __shared__ uint32_t s_cnt[3];
    __shared__ uint32_t s_b0;
    const int rep = blockIdx.x;
    const ClGeo& g = a.g;
    const uint32_t N = g.N;
    const uint32_t nw = (N + 31) / 32;
    int8_t* Sg = a.spins + (int64_t)rep * N;
    uint32_t* Sb = wl;
    uint16_t* Q = (uint16_t*)(wl + nw);
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    const int z = 2 * cl_axes(g);
    const uint32_t lane = threadIdx.x & 63u;
    // 64 sites per wave-step: the ballot of "spin is +1" is two bit words
    for (uint32_t base = (threadIdx.x >> 6) * 64u; base < N; base += blockDim.x) {
        const uint32_t i = base + lane;
        const uint64_t m = __ballot(i < N && Sg[i] > 0);
        if (lane == 0) {
            Sb[base >> 5] = (uint32_t)m;
            if ((base >> 5) + 1 < nw) Sb[(base >> 5) + 1] = (uint32_t)(m >> 32);
        }
    }
    long long total = 0;
    for (int f = 0; f < count; ++f) {
        const uint64_t sweep = a.sweeps[rep] + a.sweep_add + (uint64_t)f;
        __syncthreads();  // the previous cluster's last counter reads are done
        if (threadIdx.x == 0) {
            const u32x4 r = philox4x32_10(
                u32x4{0u, (uint32_t)sweep, ctr_z(sweep, 2, TAG_CLUSTER), sid}, k0, k1);
            const uint32_t seed = (uint32_t)(((uint64_t)r.x * N) >> 32);
            const uint32_t bit = 1u << (seed & 31u);
            s_b0 = (Sb[seed >> 5] & bit) ? 1u : 0u;
            Sb[seed >> 5] ^= bit;
            Q[0] = (uint16_t)seed;
            s_cnt[0] = 0;
        }
        __syncthreads();
        const uint32_t b0 = s_b0;
        const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
        uint32_t lo = 0, hi = 1;
        for (uint32_t lev = 0; lo < hi; ++lev) {
            const uint32_t cur = lev % 3u;
            if (threadIdx.x == 0) s_cnt[(lev + 1u) % 3u] = 0;
            // one (site, neighbour) pair per lane: a level's Philox draws
            // run side by side rather than z after each other in one lane
            const uint32_t npair = (hi - lo) * (uint32_t)z;
            for (uint32_t t = threadIdx.x; t < npair; t += blockDim.x) {
                const uint32_t k = Q[lo + t / (uint32_t)z];
                const int c = (int)(t % (uint32_t)z);
                {
                    const int ax = c >> 1;
                    uint32_t x[3];
                    if (!g.nbr) cl_coords(g, k, x);
                    const uint32_t j = g.nbr ? (uint32_t)g.nbr[(int64_t)k * g.zt + c]
                                             : cl_nbr(g, k, x, ax, (c & 1) == 0);
                    const uint32_t sh = j & 31u;
                    const uint32_t word = __hip_atomic_load(Sb + (j >> 5), __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (((word >> sh) & 1u) != b0) continue;
                    const uint32_t b = (uint32_t)ax * N + ((c & 1) ? j : k);
                    const u32x4 r = philox4x32_10(u32x4{b >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
                    const uint32_t w = b & 3;
                    const uint32_t u = w == 0 ? r.x : (w == 1 ? r.y : (w == 2 ? r.z : r.w));
                    if ((uint64_t)u >= T) continue;
                    const uint32_t bit = 1u << sh;
                    const uint32_t old = b0 ? atomicAnd(Sb + (j >> 5), ~bit)
                                            : atomicOr(Sb + (j >> 5), bit);
                    if (((old >> sh) & 1u) != b0) continue;  // another lane added it
                    Q[hi + atomicAdd(&s_cnt[cur], 1u)] = (uint16_t)j;
                }
            }
            __syncthreads();
            lo = hi;
            hi += s_cnt[cur];
        }
        total += hi;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < N; i += blockDim.x)
        Sg[i] = ((Sb[i >> 5] >> (i & 31u)) & 1u) ? (int8_t)1 : (int8_t)-1;
    if (threadIdx.x == 0) a.wsize[rep] += total;
}

__global__ void cl_advance(uint64_t* sweeps, int R, uint64_t n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < R) sweeps[t] += n;
}

// aux[rep][meas][4] = (accepted, wolff size, sum |C|^2, clusters); e/m record.
__global__ void cl_record(int R, int64_t n_meas, const int64_t* meas_d, const int32_t* e_run,
                          const int32_t* m_run, int32_t* e_out, int32_t* m_out,
                          const int64_t* acc, const int64_t* wsize, const int64_t* stats,
                          int have_sw, int64_t* aux)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= R) return;
    const int64_t o = (int64_t)rep * n_meas + *meas_d;
    e_out[o] = e_run[rep];
    m_out[o] = m_run[rep];
    aux[4 * o + 0] = acc[rep];
    aux[4 * o + 1] = wsize[rep];
    aux[4 * o + 2] = have_sw ? stats[2 * rep] : 0;
    aux[4 * o + 3] = have_sw ? stats[2 * rep + 1] : 0;
}

__global__ void cl_add_acc(int R, const int32_t* row_acc, int64_t* acc)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep < R) acc[rep] += row_acc[rep];
}

__global__ void cl_zero64(int64_t* p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// The measurement index lives on the device, so that every measurement
// launches identical kernels (and can be replayed from one captured graph).
__global__ void cl_set64(int64_t* p, int64_t v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ void cl_next(int64_t* p)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p += 1;
}

// ---------------------------------------------------------------------------
// Schedules made only of cluster rows on lattices of at most CL_LDS_SITES
// sites: one workgroup per replica runs every update and measurement of the
// call with the spins, parent / size (Swendsen-Wang) and queue / mark (Wolff)
// arrays in LDS, in one launch.  It draws the same Philox words as the kernels
// above and builds the same components (a component does not depend on the
// order its bonds are joined in, and its root is its smallest site either
// way), so the chains are identical to the multi-launch path.
constexpr int CL_LDS_SITES = 8192;
constexpr int CL_LDS_THREADS = 256;
int64_t g_cluster_lds_max_sites = CL_LDS_SITES;  // 0: always the multi-launch path

__device__ __forceinline__ uint32_t lds_ld(const int32_t* p)
{
    return (uint32_t)__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_st(int32_t* p, uint32_t v)
{
    __hip_atomic_store(p, (int32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t lds_find(int32_t* P, uint32_t x)
{
    while (true) {
        const uint32_t p = lds_ld(P + x);
        if (p == x) return x;
        const uint32_t gp = lds_ld(P + p);
        if (gp != p) lds_st(P + x, gp);
        x = gp;
    }
}

__device__ __forceinline__ void lds_unite(int32_t* P, uint32_t i, uint32_t j)
{
    while (true) {
        i = lds_find(P, i);
        j = lds_find(P, j);
        if (i == j) return;
        if (i > j) {
            const uint32_t t = i;
            i = j;
            j = t;
        }
        if (atomicCAS(P + j, (int32_t)j, (int32_t)i) == (int32_t)j) return;
    }
}

// Swendsen-Wang bonds of square lattices whose sides are multiples of
// SW_TILE, in two steps that replace sw_union.  sw_tile joins the bonds inside
// one SW_TILE x SW_TILE tile with a union-find over the tile in LDS and writes
// each site's parent as the global index of its tile component's smallest
// site; sw_edges then unites, in HBM, only the bonds that cross a tile edge
// (x or y = SW_TILE - 1 mod SW_TILE, the periodic wrap included).  A bond is
// open on the same Philox word, spins and threshold as in sw_union, local
// order is global order inside a tile, and both steps hook a larger root below
// a smaller one, so the components and their smallest-site roots are those of
// sw_union.
constexpr uint32_t SW_TILE = 64;

__host__ __device__ inline bool sw_tiled(const ClGeo& g)
{
    return !g.nbr && g.d == 2 && g.L[0] % SW_TILE == 0 && g.L[1] % SW_TILE == 0;
}

__global__ __launch_bounds__(256) void sw_tile(ClArgs a)
{
    __shared__ int8_t s_sp[SW_TILE * SW_TILE];
    __shared__ int32_t s_par[SW_TILE * SW_TILE];
    __shared__ int32_t s_sz[SW_TILE * SW_TILE];
    const int rep = blockIdx.y;
    const ClGeo& g = a.g;
    const uint32_t N = g.N, L0 = g.L[0];
    const uint32_t ntx = L0 / SW_TILE;
    const uint32_t x0 = (blockIdx.x % ntx) * SW_TILE, y0 = (blockIdx.x / ntx) * SW_TILE;
    const int8_t* S = a.spins + (int64_t)rep * N;
    int32_t* P = a.parent + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
    constexpr uint32_t NT = SW_TILE * SW_TILE;
    for (uint32_t l = threadIdx.x; l < NT; l += blockDim.x) {
        s_sp[l] = S[(y0 + l / SW_TILE) * L0 + x0 + l % SW_TILE];
        s_par[l] = (int32_t)l;
        s_sz[l] = 0;
    }
    __syncthreads();
    // groups of four bonds of one axis at sites lx0 .. lx0 + 3 of a tile row:
    // a + b = N * ax + i with i = 0 mod 4 (N and L0 are multiples of 4), so
    // the four are the words of one Philox block, as in sw_union
    constexpr uint32_t GPR = SW_TILE / 4;
    for (uint32_t gi = threadIdx.x; gi < 2 * SW_TILE * GPR; gi += blockDim.x) {
        const uint32_t ax = gi / (SW_TILE * GPR);
        const uint32_t rem = gi % (SW_TILE * GPR);
        const uint32_t ly = rem / GPR, lx0 = (rem % GPR) * 4;
        if (ax == 1 && ly == SW_TILE - 1) continue;  // crosses the tile edge
        const uint32_t b0 = ax * N + (y0 + ly) * L0 + x0 + lx0;
        const u32x4 r = philox4x32_10(u32x4{b0 >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t lx = lx0 + (uint32_t)t;
            if (ax == 0 && lx == SW_TILE - 1) continue;  // crosses the tile edge
            if ((uint64_t)w[t] >= T) continue;
            const uint32_t li = ly * SW_TILE + lx;
            const uint32_t lj = ax == 0 ? li + 1 : li + SW_TILE;
            if (s_sp[li] != s_sp[lj]) continue;
            lds_unite(s_par, li, lj);
        }
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < NT; l += blockDim.x) {
        const uint32_t r = lds_find(s_par, l);
        P[(y0 + l / SW_TILE) * L0 + x0 + l % SW_TILE] =
            (int32_t)((y0 + r / SW_TILE) * L0 + x0 + r % SW_TILE);
        atomicAdd(&s_sz[r], 1);
    }
    __syncthreads();
    // each tile component's size at its smallest site (size[] is zero elsewhere)
    for (uint32_t l = threadIdx.x; l < NT; l += blockDim.x)
        if (s_sz[l]) a.size[(int64_t)rep * N + (y0 + l / SW_TILE) * L0 + x0 + l % SW_TILE] = s_sz[l];
}

// The bonds that cross a tile edge: L1 * L0 / SW_TILE of axis 0, then
// L0 * L1 / SW_TILE of axis 1.
__global__ __launch_bounds__(256) void sw_edges(ClArgs a)
{
    const int rep = blockIdx.y;
    const ClGeo& g = a.g;
    const uint32_t N = g.N, L0 = g.L[0], L1 = g.L[1];
    const uint32_t ntx = L0 / SW_TILE, nty = L1 / SW_TILE;
    const uint32_t E0 = L1 * ntx, E = E0 + L0 * nty;
    const int8_t* S = a.spins + (int64_t)rep * N;
    int32_t* P = a.parent + (int64_t)rep * N;
    const uint64_t sweep = a.sweeps[rep] + a.sweep_add;
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        uint32_t ax, i;
        if (e < E0) {
            ax = 0;
            i = (e / ntx) * L0 + (e % ntx) * SW_TILE + SW_TILE - 1;
        } else {
            const uint32_t e1 = e - E0;
            ax = 1;
            i = ((e1 / L0) * SW_TILE + SW_TILE - 1) * L0 + e1 % L0;
        }
        const uint32_t b = ax * N + i;
        const u32x4 r = philox4x32_10(u32x4{b >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
        const uint32_t w = b & 3;
        const uint32_t u = w == 0 ? r.x : (w == 1 ? r.y : (w == 2 ? r.z : r.w));
        if ((uint64_t)u >= T) continue;
        const uint32_t j = cl_nbr_col(g, i, 2 * (int)ax);
        if (S[i] != S[j]) continue;
        cl_unite(P, i, j);
    }
}

// sw_flatten after sw_tile / sw_edges: size[] holds each tile component's
// size at its smallest site, so a site that is such a component's root but
// not its cluster's adds that size to the cluster's root, and every other
// site only stores its root.  The root's own entry is read by no other lane.
__global__ __launch_bounds__(256) void sw_flatten_tiled(ClArgs a)
{
    const int rep = blockIdx.y;
    const uint32_t N = a.g.N;
    int32_t* P = a.parent + (int64_t)rep * N;
    int32_t* Z = a.size + (int64_t)rep * N;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint32_t r = cl_find_ro(P, i);
        st_relaxed(P + i, r);
        if (r != i) {
            const int32_t z = Z[i];
            if (z) atomicAdd(Z + r, z);
        }
    }
}

// Sum of v over the workgroup, returned to every lane (uses red[0..3]).
__device__ __forceinline__ long long cl_block_sum(long long v, long long* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    return t;
}

__global__ __launch_bounds__(CL_LDS_THREADS) void cl_lds_run(ClArgs a, SeqRows rows, int64_t n_meas,
                                                             uint64_t* sweeps, int32_t* e_run,
                                                             int32_t* mc_run, int32_t* e_out,
                                                             int32_t* mc_out, int64_t* aux_out)
{
    extern __shared__ __align__(16) unsigned char cl_smem[];
    __shared__ long long red[CL_LDS_THREADS / 64];
    __shared__ uint32_t s_head, s_tail;
    __shared__ int s_a, s_b;
    __shared__ int32_t s_cnt[128];
    const ClGeo g = a.g;
    const uint32_t N = g.N;
    const int axes = cl_axes(g), z = 2 * axes, q = a.q;
    int32_t* A = (int32_t*)cl_smem;   // parent (SW) or queue (Wolff)
    int32_t* B = A + N;               // size (SW) or mark (Wolff)
    int8_t* S = (int8_t*)(B + N);     // spins
    const int rep = blockIdx.x;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    int8_t* Sg = a.spins + (int64_t)rep * N;
    for (uint32_t i = tid; i < N; i += bs) S[i] = Sg[i];
    const uint32_t k0 = a.keys[2 * rep], k1 = a.keys[2 * rep + 1], sid = a.sids[rep];
    const uint64_t T = a.t_add[rep];
    uint64_t sweep = a.sweeps[rep];
    long long e = 0;
    __syncthreads();
    for (int64_t meas = 0; meas < n_meas; ++meas) {
        long long wsize = 0, c2 = 0, ncl = 0;
        for (int r = 0; r < rows.n; ++r) {
            if (rows.code[r] == CODE_WOLFF) {
                for (uint32_t i = tid; i < N; i += bs) B[i] = -1;
                __syncthreads();
            }
            for (int f = 0; f < rows.count[r]; ++f, ++sweep) {
                if (rows.code[r] == CODE_SWENDSEN_WANG) {
                    for (uint32_t i = tid; i < N; i += bs) {
                        A[i] = (int32_t)i;
                        B[i] = 0;
                    }
                    __syncthreads();
                    const uint32_t nb = (uint32_t)axes * N;
                    const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
                    for (uint32_t blk = tid; 4 * blk < nb; blk += bs) {
                        const u32x4 w4 = philox4x32_10(u32x4{blk, (uint32_t)sweep, zc, sid}, k0, k1);
                        const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const uint32_t b = 4 * blk + t;
                            if (b >= nb) break;
                            if ((uint64_t)w[t] >= T) continue;
                            int ax = 0;
                            uint32_t i = b;
                            while (i >= N) {
                                i -= N;
                                ++ax;
                            }
                            const uint32_t j = cl_nbr_col(g, i, 2 * ax);
                            if (S[i] != S[j]) continue;
                            lds_unite(A, i, j);
                        }
                    }
                    __syncthreads();
                    for (uint32_t i = tid; i < N; i += bs) {
                        uint32_t x = i;
                        while (true) {  // read-only find: every entry written here is a root
                            const uint32_t p = lds_ld(A + x);
                            if (p == x) break;
                            x = p;
                        }
                        lds_st(A + i, x);
                        atomicAdd(B + x, 1);
                    }
                    __syncthreads();
                    // each cluster's new state is drawn once, at its root (whose size
                    // entry is read first), then every site takes it from there
                    long long tc2 = 0, tn = 0;
                    for (uint32_t i = tid; i < N; i += bs) {
                        if ((uint32_t)A[i] != i) continue;
                        tc2 += (long long)B[i] * B[i];
                        tn += 1;
                        const uint32_t x = cl_word(k0, k1, sid, sweep, 1, i);
                        B[i] = q == 0 ? (int32_t)((x >> 31) == 0)
                                      : (int32_t)(((uint64_t)x * (uint32_t)q) >> 32);
                    }
                    __syncthreads();
                    for (uint32_t i = tid; i < N; i += bs) {
                        const int32_t d = B[A[i]];
                        if (q == 0) {
                            if (d) S[i] = (int8_t)(-S[i]);
                        } else {
                            S[i] = (int8_t)d;
                        }
                    }
                    c2 = cl_block_sum(tc2, red);
                    ncl = cl_block_sum(tn, red);
                    __syncthreads();
                } else {  // one Wolff cluster
                    if (tid == 0) {
                        const u32x4 w4 = philox4x32_10(
                            u32x4{0u, (uint32_t)sweep, ctr_z(sweep, 2, TAG_CLUSTER), sid}, k0, k1);
                        const uint32_t seed = (uint32_t)(((uint64_t)w4.x * N) >> 32);
                        const int av = S[seed];
                        int bv;
                        if (q == 0) {
                            bv = -av;
                        } else {
                            bv = (int)(((uint64_t)w4.y * (uint32_t)(q - 1)) >> 32);
                            if (bv >= av) ++bv;
                        }
                        S[seed] = (int8_t)bv;
                        B[seed] = f;
                        A[0] = (int32_t)seed;
                        s_a = av;
                        s_b = bv;
                        s_head = 0;
                        s_tail = 1;
                    }
                    __syncthreads();
                    const int sa0 = s_a, sb1 = s_b;
                    const uint32_t zc = ctr_z(sweep, 0, TAG_CLUSTER);
                    while (true) {
                        const uint32_t lo = s_head, hi = s_tail;
                        // every lane has this level's bounds before any lane appends
                        __syncthreads();
                        if (lo == hi) break;
                        // one (site, neighbour) pair per lane
                        const uint32_t npair = (hi - lo) * (uint32_t)z;
                        for (uint32_t t = (uint32_t)tid; t < npair; t += (uint32_t)bs) {
                            const uint32_t k = (uint32_t)A[lo + t / (uint32_t)z];
                            const int c = (int)(t % (uint32_t)z);
                            {
                                const uint32_t j = cl_nbr_col(g, k, c);
                                if (S[j] != sa0) continue;
                                const uint32_t b = (uint32_t)(c >> 1) * N + ((c & 1) ? j : k);
                                const u32x4 w4 = philox4x32_10(u32x4{b >> 2, (uint32_t)sweep, zc, sid}, k0, k1);
                                const uint32_t wi = b & 3;
                                const uint32_t u = wi == 0 ? w4.x : (wi == 1 ? w4.y : (wi == 2 ? w4.z : w4.w));
                                if ((uint64_t)u >= T) continue;
                                if (atomicExch(B + j, f) == f) continue;
                                S[j] = (int8_t)sb1;
                                A[atomicAdd(&s_tail, 1u)] = (int32_t)j;
                            }
                        }
                        __syncthreads();
                        if (tid == 0) s_head = hi;
                        __syncthreads();
                    }
                    wsize += s_tail;
                    __syncthreads();
                }
            }
        }
        // recount the energy and magnetisation / occupation counts
        if (q)
            for (int k = tid; k < q; k += bs) s_cnt[k] = 0;
        __syncthreads();
        long long te = 0, tm = 0;
        for (uint32_t i = tid; i < N; i += bs) {
            const int si = S[i];
            for (int ax = 0; ax < axes; ++ax) {
                const int sj = S[cl_nbr_col(g, i, 2 * ax)];
                te -= q ? (long long)(si == sj) : (long long)(si * sj);
            }
            if (q) atomicAdd(s_cnt + si, 1);
            else tm += si;
        }
        e = cl_block_sum(te, red);
        const long long m = q ? 0 : cl_block_sum(tm, red);
        __syncthreads();
        const int64_t o = (int64_t)rep * n_meas + meas;
        if (tid == 0) {
            e_out[o] = (int32_t)e;
            if (!q) mc_out[o] = (int32_t)m;
            aux_out[4 * o + 0] = 0;
            aux_out[4 * o + 1] = wsize;
            aux_out[4 * o + 2] = c2;
            aux_out[4 * o + 3] = ncl;
            if (!q) mc_run[rep] = (int32_t)m;
        }
        if (q)
            for (int k = tid; k < q; k += bs) {
                mc_out[o * q + k] = s_cnt[k];
                mc_run[(int64_t)rep * q + k] = s_cnt[k];
            }
        __syncthreads();
    }
    for (uint32_t i = tid; i < N; i += bs) Sg[i] = S[i];
    if (tid == 0) {
        sweeps[rep] = sweep;
        e_run[rep] = (int32_t)e;
    }
}

// Launch cl_lds_run when the schedule has only cluster rows and the lattice
// fits; returns false (nothing launched) otherwise.
inline bool cl_lds_try(const ClArgs& a, const int64_t* schedule, int n_rows, int64_t n_meas,
                       uint64_t* sweeps, int32_t* e_run, int32_t* mc_run, int32_t* e_out,
                       int32_t* mc_out, int64_t* aux_out, hipStream_t st, hipError_t* err)
{
    if ((int64_t)a.g.N > std::min<int64_t>(g_cluster_lds_max_sites, CL_LDS_SITES)) return false;
    SeqRows rows{};
    rows.n = n_rows;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r];
        if (code != CODE_WOLFF && code != CODE_SWENDSEN_WANG) return false;
        rows.code[r] = (int32_t)code;
        rows.count[r] = (int32_t)schedule[2 * r + 1];
    }
    const size_t lds = (size_t)a.g.N * 9 + 16;
    *err = hipSuccess;
    if (lds > 64 * 1024)
        *err = hipFuncSetAttribute((const void*)cl_lds_run,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (*err == hipSuccess) {  // 256 lanes: 64 or 128 measured slower for SW on 16^2 and 32^2
        hipLaunchKernelGGL(cl_lds_run, dim3(a.R), dim3(CL_LDS_THREADS), lds, st, a, rows, n_meas,
                           sweeps, e_run, mc_run, e_out, mc_out, aux_out);
        *err = hipGetLastError();
    }
    return true;
}

inline uint64_t cl_threshold(double p)
{
    if (!(p < 1.0)) return 1ull << 32;
    if (!(p > 0.0)) return 0;
    return (uint64_t)std::ceil(p * 4294967296.0);
}

__global__ void cl_record_potts(int R, int q, int64_t n_meas, const int64_t* meas_d,
                                const int32_t* e_run, const int32_t* counts_run, int32_t* e_out,
                                int32_t* counts_out, const int64_t* acc, const int64_t* wsize,
                                const int64_t* stats, int have_sw, int64_t* aux)
{
    const int rep = blockIdx.x * blockDim.x + threadIdx.x;
    if (rep >= R) return;
    const int64_t o = (int64_t)rep * n_meas + *meas_d;
    e_out[o] = e_run[rep];
    for (int k = 0; k < q; ++k) counts_out[o * q + k] = counts_run[(int64_t)rep * q + k];
    aux[4 * o + 0] = acc[rep];
    aux[4 * o + 1] = wsize[rep];
    aux[4 * o + 2] = have_sw ? stats[2 * rep] : 0;
    aux[4 * o + 3] = have_sw ? stats[2 * rep + 1] : 0;
}

struct ClLayout {
    int64_t tab, parent, size, queue, mark, stats, wsize, acc, tadd, rowacc, rowe, rowm, meas, total;
};

// Workspace: the local kernels' tables (tab_bytes), then the cluster arrays;
// rowm holds row_m_words int32 per replica (M for Ising, q counts for Potts).
inline ClLayout cl_layout(int64_t N, int32_t n_rep, int64_t tab_bytes, int row_m_words)
{
    auto up = [](int64_t x) { return (x + 255) / 256 * 256; };
    ClLayout l{};
    int64_t o = 0;
    l.tab = o;    o += up(tab_bytes);
    l.parent = o; o += up(4 * N * n_rep);
    l.size = o;   o += up(4 * N * n_rep);
    l.queue = o;  o += up(4 * N * n_rep);
    l.mark = o;   o += up(4 * N * n_rep);
    l.stats = o;  o += up(16 * (int64_t)n_rep);
    l.wsize = o;  o += up(8 * (int64_t)n_rep);
    l.acc = o;    o += up(8 * (int64_t)n_rep);
    l.tadd = o;   o += up(8 * (int64_t)n_rep);
    l.rowacc = o; o += up(4 * (int64_t)n_rep);
    l.rowe = o;   o += up(4 * (int64_t)n_rep);
    l.rowm = o;   o += up(4 * (int64_t)n_rep * row_m_words);
    l.meas = o;   o += up(8);
    l.total = o;
    return l;
}

// Validate the geometry and schedule shared by both entry points.
inline int cl_check(int32_t ndim, const int32_t* dims, int32_t n_rep, int64_t n_meas,
                    const int64_t* schedule, int32_t n_rows, int64_t* N_out, bool* any_cluster,
                    bool* any_sw)
{
    CM_REQUIRE(ndim >= 1 && ndim <= 3, "lattices of 1..3 dimensions");
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) {
        CM_REQUIRE(dims[a] >= 2 && dims[a] % 2 == 0, "extents must be even and >= 2");
        N *= dims[a];
    }
    CM_REQUIRE(N * ndim < ((int64_t)1 << 31), "lattice too large for cluster updates");
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    CM_REQUIRE(n_meas >= 0, "n_meas must not be negative");
    CM_REQUIRE(n_rows >= 1 && n_rows <= 8, "schedule must have 1..8 rows");
    *any_cluster = *any_sw = false;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r], count = schedule[2 * r + 1];
        CM_REQUIRE(code == CODE_METROPOLIS || code == CODE_HEATBATH || code == CODE_WOLFF ||
                       code == CODE_SWENDSEN_WANG,
                   "update code %lld has no GPU kernel for this model", (long long)code);
        CM_REQUIRE(count >= 1 && count <= (1 << 24), "repetition count out of range");
        *any_cluster |= code == CODE_WOLFF || code == CODE_SWENDSEN_WANG;
        *any_sw |= code == CODE_SWENDSEN_WANG;
    }
    *N_out = N;
    return CM_OK;
}

inline void cl_args(ClArgs& a, char* ws, const ClLayout& lay, void* spins, int32_t ndim,
                    const int32_t* dims, int64_t N, int32_t n_rep, int q, const uint32_t* keys,
                    const uint32_t* sids, uint64_t* sweeps)
{
    a.spins = (int8_t*)spins;
    a.parent = (int32_t*)(ws + lay.parent);
    a.size = (int32_t*)(ws + lay.size);
    a.queue = (int32_t*)(ws + lay.queue);
    a.mark = (int32_t*)(ws + lay.mark);
    a.stats = (int64_t*)(ws + lay.stats);
    a.wsize = (int64_t*)(ws + lay.wsize);
    a.t_add = (const uint64_t*)(ws + lay.tadd);
    a.keys = keys;
    a.sids = sids;
    a.sweeps = sweeps;
    a.g.d = ndim;
    for (int k = 0; k < 3; ++k) a.g.L[k] = k < ndim ? (uint32_t)dims[k] : 1u;
    a.g.N = (uint32_t)N;
    a.g.div0 = make_fastdiv(a.g.L[0]);
    a.g.div1 = make_fastdiv(a.g.L[1]);
    a.g.nbr = nullptr;
    a.g.zt = 2 * ndim;
    a.R = n_rep;
    a.q = q;
}

// `count` Swendsen-Wang updates or Wolff clusters of every replica, then the
// device sweep counters advance by `count`.
inline void cl_cluster_row(ClArgs a, int64_t code, int64_t count, uint64_t* sweeps,
                           hipStream_t st)
{
    const int64_t N = a.g.N;
    const unsigned site_chunks = (unsigned)std::min<int64_t>((N + 255) / 256, 1024);
    const unsigned bond_chunks = (unsigned)std::min<int64_t>((cl_axes(a.g) * N / 4 + 256) / 256, 1024);
    if (code == CODE_SWENDSEN_WANG) {
        const bool tiled = sw_tiled(a.g);
        const unsigned flip_chunks = (unsigned)std::min<int64_t>((N + 4095) / 4096, 1024);
        const unsigned n_tiles = tiled ? (a.g.L[0] / SW_TILE) * (a.g.L[1] / SW_TILE) : 0;
        const unsigned edge_chunks =
            tiled ? (unsigned)std::min<int64_t>(
                        ((int64_t)a.g.L[1] * (a.g.L[0] / SW_TILE) +
                         (int64_t)a.g.L[0] * (a.g.L[1] / SW_TILE) + 255) / 256, 1024)
                  : 0;
        for (int64_t k = 0; k < count; ++k) {
            a.sweep_add = (uint64_t)k;
            hipLaunchKernelGGL(sw_init, dim3(site_chunks, a.R), dim3(256), 0, st, a);
            if (tiled) {
                hipLaunchKernelGGL(sw_tile, dim3(n_tiles, a.R), dim3(256), 0, st, a);
                hipLaunchKernelGGL(sw_edges, dim3(edge_chunks, a.R), dim3(256), 0, st, a);
                hipLaunchKernelGGL(sw_flatten_tiled, dim3(site_chunks, a.R), dim3(256), 0, st, a);
            } else {
                hipLaunchKernelGGL(sw_union, dim3(bond_chunks, a.R), dim3(256), 0, st, a);
                hipLaunchKernelGGL(sw_flatten, dim3(site_chunks, a.R), dim3(256), 0, st, a);
            }
            // 16 sites per lane: each workgroup ends with one 64-bit atomic
            // pair on its replica's statistics, and these serialise
            hipLaunchKernelGGL(sw_flip, dim3(flip_chunks, a.R), dim3(256), 0, st, a);
        }
    } else {
        a.sweep_add = 0;
        if (a.q == 0 && N <= CL_WOLFF_LDS_SITES) {
            const size_t lds = (size_t)(N + 31) / 32 * 4 + (size_t)(N + 1) / 2 * 4;
            if (lds > 64 * 1024)
                (void)hipFuncSetAttribute((const void*)wolff_ising_bits,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(wolff_ising_bits, dim3(a.R), dim3(256), lds, st, a, (int)count);
        } else if (N <= CL_WOLFF_LDS_SITES) {
            const size_t lds = (size_t)(N + 3) / 4 * 4 + (size_t)(N + 31) / 32 * 4;
            // a launch beyond the attribute fails, and the callers' launch
            // check reports it
            if (lds > 64 * 1024)
                (void)hipFuncSetAttribute((const void*)wolff_run_lds,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(wolff_run_lds, dim3(a.R), dim3(256), lds, st, a, (int)count);
        } else {
            const int64_t nm = N * a.R;
            hipLaunchKernelGGL(cl_fill, dim3((unsigned)std::min<int64_t>((nm + 255) / 256, 8192)),
                               dim3(256), 0, st, a.mark, nm, -1);
            hipLaunchKernelGGL(wolff_run, dim3(a.R), dim3(256), 0, st, a, (int)count);
        }
    }
    hipLaunchKernelGGL(cl_advance, dim3((a.R + 255) / 256), dim3(256), 0, st, sweeps, a.R,
                       (uint64_t)count);
}

// ---------------------------------------------------------------------------
// Measurement loops as graphs.  A measurement of a cluster schedule is a dozen
// small launches (fills, the Swendsen-Wang passes or the Wolff search, local
// sweeps, a recount, the record), so on small lattices the loop is bound by
// launch cost.  When every launch of a measurement is a plain kernel launch
// with the same arguments each time (the measurement index and the sweep
// counters live on the device), one measurement is captured into a hipGraph on
// a private stream and the graph is replayed n times on the caller's stream.

// Measurements from which a loop is replayed as a graph (0: never).  Off by
// default: on MI355X a replayed measurement took 0.89-1.02x the eager time
// (SW and Wolff on 16^2-64^2, 1-64 replicas), because each node still costs a
// kernel dispatch on the device; the host launch cost it removes is hidden.
int64_t g_graph_min_meas = 0;

// defined with the tempering kernels
extern int64_t g_ea_fused, g_ea_order;

namespace {
struct Retired {
    hipGraphExec_t exec;
    hipEvent_t done;  // recorded after the graph's last replay
};
thread_local std::vector<Retired> g_retired;
thread_local hipStream_t g_capture[64] = {};

void release_finished_graphs()
{
    size_t keep = 0;
    for (Retired& r : g_retired) {
        if (hipEventQuery(r.done) == hipSuccess) {
            (void)hipGraphExecDestroy(r.exec);
            (void)hipEventDestroy(r.done);
        } else {
            g_retired[keep++] = r;
        }
    }
    g_retired.resize(keep);
}
}  // namespace

// body(stream, eager) enqueues one measurement.  Runs it n times on st: eagerly
// (eager = true), or, when `capturable` and n reaches g_graph_min_meas,
// captured once (eager = false) and replayed n times.  *graphed says which.
template <class F>
int repeat_measurements(int64_t n, bool capturable, hipStream_t st, bool* graphed, F&& body)
{
    *graphed = false;
    release_finished_graphs();
    if (!capturable || g_graph_min_meas <= 0 || n < g_graph_min_meas) {
        for (int64_t i = 0; i < n; ++i)
            if (int rc = body(st, true)) return rc;
        return CM_OK;
    }
    int dev = 0;
    CM_HIP_CHECK(hipGetDevice(&dev));
    CM_REQUIRE(dev >= 0 && dev < 64, "device %d out of range", dev);
    hipStream_t& cap = g_capture[dev];
    if (cap == nullptr) CM_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    CM_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    const int rc = body(cap, false);
    hipGraph_t graph = nullptr;
    const hipError_t ended = hipStreamEndCapture(cap, &graph);
    if (rc != CM_OK) {
        if (ended == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    CM_HIP_CHECK(ended);
    hipGraphExec_t exec = nullptr;
    const hipError_t made = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CM_HIP_CHECK(made);
    hipError_t e = hipSuccess;
    for (int64_t i = 0; i < n && e == hipSuccess; ++i) e = hipGraphLaunch(exec, st);
    hipEvent_t done = nullptr;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(done, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        if (done) (void)hipEventDestroy(done);
        CM_HIP_CHECK(e);
    }
    g_retired.push_back({exec, done});
    *graphed = true;
    return CM_OK;
}

// Per-chain tables of the sequential kernels (seq_core.hpp) and the cluster
// bond thresholds, from the reference's FP64 formulas: Ising
// critmc/kernels/ising.py:47-65 and p_add = 1 - e^{-2 beta} (:201); Potts
// e^{-beta d}, e^{beta n} and p_add = 1 - e^{-beta} (kernels/potts.py:194-200).
inline void seq_tables(int q, int z, int n_rep, const double* betas, double h,
                       std::vector<uint64_t>& tab, std::vector<uint64_t>& tadd)
{
    tab.assign((size_t)n_rep * SEQ_TAB, 0);
    tadd.resize(n_rep);
    for (int r = 0; r < n_rep; ++r) {
        uint64_t* t = &tab[(size_t)r * SEQ_TAB];
        const double b = betas[r];
        if (q == 0) {
            for (int al = -z; al <= z; ++al) {
                for (int up = 0; up < 2; ++up) {
                    const double de = 2.0 * al + 2.0 * h * (2 * up - 1);
                    t[2 * (al + z) + up] = cl_threshold(de <= 0.0 ? 1.0 : std::exp(-b * de));
                }
                t[SEQ_HB + al + z] = cl_threshold(1.0 / (1.0 + std::exp(-2.0 * b * (al + h))));
            }
            tadd[r] = cl_threshold(1.0 - std::exp(-2.0 * b));
        } else {
            double* w = (double*)(t + SEQ_W);
            for (int d = 0; d <= z; ++d) {
                t[d] = cl_threshold(std::exp(-b * d));
                w[d] = std::exp(b * d);
            }
            tadd[r] = cl_threshold(1.0 - std::exp(-b));
        }
    }
}

// Schedules of any lattice given by its neighbour table: local rows run as
// sequential chains (seq_run), cluster rows through the same Wolff /
// Swendsen-Wang kernels as the hypercubic path with neighbours from the table.
inline int table_run(int q, void* spins, int64_t N, int32_t z, const int32_t* nbr, int32_t n_rep,
                     const double* betas, double h, const uint32_t* keys, const uint32_t* sids,
                     uint64_t* sweeps, const int64_t* schedule, int32_t n_rows, int64_t n_meas,
                     int32_t* e_run, int32_t* mc_run, int32_t* e_out, int32_t* mc_out,
                     int64_t* aux_out, void* workspace, int64_t workspace_bytes, hipStream_t st)
{
    CM_REQUIRE(N >= 2 && N * z < ((int64_t)1 << 31), "lattice of %lld sites out of range",
               (long long)N);
    CM_REQUIRE(z >= 2 && z <= SEQ_ZMAX && z % 2 == 0,
               "neighbour table needs an even number of columns in 2..%d, got %d", SEQ_ZMAX, z);
    CM_REQUIRE(nbr != nullptr, "the neighbour table is required");
    CM_REQUIRE(n_rep >= 1 && n_rep <= 65535, "n_rep out of range");
    CM_REQUIRE(n_meas >= 0, "n_meas must not be negative");
    CM_REQUIRE(n_rows >= 1 && n_rows <= 8, "schedule must have 1..8 rows");
    CM_REQUIRE(std::isfinite(h), "field h must be finite");
    bool any_cluster = false, any_sw = false;
    SeqRows all{};
    all.n = n_rows;
    for (int r = 0; r < n_rows; ++r) {
        const int64_t code = schedule[2 * r], count = schedule[2 * r + 1];
        CM_REQUIRE(code >= CODE_METROPOLIS && code <= CODE_SWENDSEN_WANG,
                   "update code %lld has no GPU kernel for this model", (long long)code);
        CM_REQUIRE(count >= 1 && count <= (1 << 24), "repetition count out of range");
        any_cluster |= code == CODE_WOLFF || code == CODE_SWENDSEN_WANG;
        any_sw |= code == CODE_SWENDSEN_WANG;
        all.code[r] = (int32_t)code;
        all.count[r] = (int32_t)count;
    }
    CM_REQUIRE(!any_cluster || h == 0.0, "cluster updates require zero field h");
    const ClLayout lay = cl_layout(N, n_rep, (int64_t)n_rep * SEQ_TAB * 8, q > 0 ? q : 1);
    CM_REQUIRE(workspace_bytes >= lay.total, "workspace too small");
    for (int r = 0; r < n_rep; ++r) CM_REQUIRE(std::isfinite(betas[r]), "beta must be finite");
    if (n_meas == 0) return CM_OK;
    char* ws = (char*)workspace;
    std::vector<uint64_t> tab, tadd;
    seq_tables(q, z, n_rep, betas, h, tab, tadd);
    CM_HIP_CHECK(upload_async(ws + lay.tab, tab.data(), tab.size() * 8, st));
    CM_HIP_CHECK(upload_async(ws + lay.tadd, tadd.data(), tadd.size() * 8, st));
    SeqArgs sa{};
    sa.spins = (int8_t*)spins;
    sa.nbr = nbr;
    sa.N = (uint32_t)N;
    sa.z = z;
    sa.R = n_rep;
    sa.q = q;
    sa.tab = (const uint64_t*)(ws + lay.tab);
    sa.keys = keys;
    sa.sids = sids;
    sa.sweeps = sweeps;
    sa.e_run = e_run;
    sa.m_run = q ? nullptr : mc_run;
    sa.counts = q ? mc_run : nullptr;
    const dim3 lanes((unsigned)((n_rep + 63) / 64));
    if (!any_cluster) {  // the whole call in one launch
        sa.acc = nullptr;
        hipLaunchKernelGGL(seq_run, lanes, dim3(64), 0, st, sa, all, n_meas, e_out, mc_out, aux_out);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    }
    ClArgs ca{};
    ca.spins = (int8_t*)spins;
    ca.parent = (int32_t*)(ws + lay.parent);
    ca.size = (int32_t*)(ws + lay.size);
    ca.queue = (int32_t*)(ws + lay.queue);
    ca.mark = (int32_t*)(ws + lay.mark);
    ca.stats = (int64_t*)(ws + lay.stats);
    ca.wsize = (int64_t*)(ws + lay.wsize);
    ca.t_add = (const uint64_t*)(ws + lay.tadd);
    ca.keys = keys;
    ca.sids = sids;
    ca.sweeps = sweeps;
    ca.g.d = z / 2;
    ca.g.L[0] = ca.g.L[1] = ca.g.L[2] = 1;
    ca.g.N = (uint32_t)N;
    ca.g.div0 = make_fastdiv(1);
    ca.g.div1 = make_fastdiv(1);
    ca.g.nbr = nbr;
    ca.g.zt = z;
    ca.R = n_rep;
    ca.q = q;
    hipError_t lds_err;
    if (cl_lds_try(ca, schedule, n_rows, n_meas, sweeps, e_run, mc_run, e_out, mc_out, aux_out,
                   st, &lds_err)) {
        CM_HIP_CHECK(lds_err);
        return CM_OK;
    }
    int64_t* acc = (int64_t*)(ws + lay.acc);
    int64_t* meas_d = (int64_t*)(ws + lay.meas);
    sa.acc = acc;
    const int rb = (n_rep + 255) / 256;
    hipLaunchKernelGGL(cl_zero64, dim3((2 * n_rep + 255) / 256), dim3(256), 0, st, ca.stats, 2 * n_rep);
    hipLaunchKernelGGL(cl_set64, dim3(1), dim3(64), 0, st, meas_d, (int64_t)0);
    auto measurement = [&](hipStream_t s, bool) -> int {
        hipLaunchKernelGGL(cl_zero64, dim3(rb), dim3(256), 0, s, acc, n_rep);
        hipLaunchKernelGGL(cl_zero64, dim3(rb), dim3(256), 0, s, ca.wsize, n_rep);
        for (int r = 0; r < n_rows; ++r) {
            const int code = all.code[r];
            if (code == CODE_WOLFF || code == CODE_SWENDSEN_WANG) {
                cl_cluster_row(ca, code, all.count[r], sweeps, s);
                hipLaunchKernelGGL(seq_measure, dim3(n_rep), dim3(256), 0, s, sa);
            } else {
                SeqRows one{};
                one.n = 1;
                one.code[0] = code;
                one.count[0] = all.count[r];
                hipLaunchKernelGGL(seq_run, lanes, dim3(64), 0, s, sa, one, (int64_t)1,
                                   (int32_t*)nullptr, (int32_t*)nullptr, (int64_t*)nullptr);
            }
        }
        if (q)
            hipLaunchKernelGGL(cl_record_potts, dim3(rb), dim3(256), 0, s, n_rep, q, n_meas,
                               (const int64_t*)meas_d, e_run, mc_run, e_out, mc_out, acc,
                               ca.wsize, ca.stats, any_sw ? 1 : 0, aux_out);
        else
            hipLaunchKernelGGL(cl_record, dim3(rb), dim3(256), 0, s, n_rep, n_meas,
                               (const int64_t*)meas_d, e_run, mc_run, e_out, mc_out, acc,
                               ca.wsize, ca.stats, any_sw ? 1 : 0, aux_out);
        hipLaunchKernelGGL(cl_next, dim3(1), dim3(64), 0, s, meas_d);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    };
    bool graphed;
    return repeat_measurements(n_meas, true, st, &graphed, measurement);
}

}  // namespace cm

using namespace cm;

extern "C" {

int64_t cm_ising_cluster_workspace_bytes(int32_t ndim, const int32_t* dims, int32_t n_rep)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    return cl_layout(N, n_rep, cm_ising_workspace_bytes(n_rep), 1).total;
}

int cm_ising_cluster_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep,
                         const double* betas, double h, const uint32_t* keys,
                         const uint32_t* stream_ids, uint64_t* sweeps, uint64_t* sweeps_host,
                         const int64_t* schedule, int32_t n_rows, int64_t n_meas, int32_t* e_run,
                         int32_t* m_run, int32_t* e_out, int32_t* m_out, int64_t* aux_out,
                         void* workspace, int64_t workspace_bytes, void* stream)
{
    int64_t N;
    bool any_cluster, any_sw;
    if (int rc = cl_check(ndim, dims, n_rep, n_meas, schedule, n_rows, &N, &any_cluster, &any_sw))
        return rc;
    CM_REQUIRE(sweeps_host != nullptr, "the host sweep counters are required");
    const ClLayout lay = cl_layout(N, n_rep, cm_ising_workspace_bytes(n_rep), 1);
    CM_REQUIRE(workspace_bytes >= lay.total, "workspace too small");
    CM_REQUIRE(!any_cluster || h == 0.0, "cluster updates require zero field h");
    if (n_meas == 0) return CM_OK;
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    std::vector<uint64_t> tadd(n_rep);
    for (int r = 0; r < n_rep; ++r) {
        CM_REQUIRE(std::isfinite(betas[r]), "beta must be finite");
        tadd[r] = cl_threshold(1.0 - std::exp(-2.0 * betas[r]));
    }
    CM_HIP_CHECK(upload_async(ws + lay.tadd, tadd.data(), tadd.size() * 8, st));
    ClArgs a{};
    cl_args(a, ws, lay, spins, ndim, dims, N, n_rep, 0, keys, stream_ids, sweeps);
    int64_t* acc = (int64_t*)(ws + lay.acc);
    int32_t* row_acc = (int32_t*)(ws + lay.rowacc);
    int32_t* row_e = (int32_t*)(ws + lay.rowe);
    int32_t* row_m = (int32_t*)(ws + lay.rowm);
    const int rb = (n_rep + 255) / 256;
    // the local rows' class tables are built and uploaded once per call; the
    // record of each update code says which class loop its launches need
    int32_t prep[2][2] = {{0, 0}, {0, 0}};
    for (int row = 0; row < n_rows; ++row) {
        const int64_t code = schedule[2 * row];
        if (code != CODE_METROPOLIS && code != CODE_HEATBATH) continue;
        int32_t* rec = prep[code == CODE_HEATBATH ? 1 : 0];
        if (rec[0] != 0) continue;
        const int64_t one[2] = {code, 1};
        int rc = cm_ising_prepare(ndim, dims, n_rep, betas, h, one, 1, ws + lay.tab,
                                  lay.parent - lay.tab, rec, stream);
        if (rc != CM_OK) return rc;
    }
    int64_t* meas_d = (int64_t*)(ws + lay.meas);
    int64_t per_meas = 0;
    bool any_local = false;
    for (int row = 0; row < n_rows; ++row) {
        per_meas += schedule[2 * row + 1];
        any_local |= schedule[2 * row] == CODE_METROPOLIS || schedule[2 * row] == CODE_HEATBATH;
    }
    hipError_t lds_err;
    if (cl_lds_try(a, schedule, n_rows, n_meas, sweeps, e_run, m_run, e_out, m_out, aux_out, st,
                   &lds_err)) {
        CM_HIP_CHECK(lds_err);
        for (int r = 0; r < n_rep; ++r) sweeps_host[r] += (uint64_t)(per_meas * n_meas);
        return CM_OK;
    }
    hipLaunchKernelGGL(cl_zero64, dim3((2 * n_rep + 255) / 256), dim3(256), 0, st, a.stats, 2 * n_rep);
    hipLaunchKernelGGL(cl_set64, dim3(1), dim3(64), 0, st, meas_d, (int64_t)0);
    // eager: the host counters advance row by row, mirroring the device ones
    auto measurement = [&](hipStream_t s, bool eager) -> int {
        hipLaunchKernelGGL(cl_zero64, dim3(rb), dim3(256), 0, s, acc, n_rep);
        hipLaunchKernelGGL(cl_zero64, dim3(rb), dim3(256), 0, s, a.wsize, n_rep);
        for (int row = 0; row < n_rows; ++row) {
            const int64_t code = schedule[2 * row], count = schedule[2 * row + 1];
            if (code == CODE_METROPOLIS || code == CODE_HEATBATH) {
                const int64_t one[2] = {code, count};
                int rc = cm_ising_run(spins, 0, ndim, dims, n_rep, betas, h, keys, stream_ids,
                                      sweeps, sweeps_host, one, 1, 1, e_run, m_run, row_e, row_m,
                                      row_acc, ws + lay.tab, lay.parent - lay.tab, 0, 0, 0,
                                      nullptr, prep[code == CODE_HEATBATH ? 1 : 0], s);
                if (rc != CM_OK) return rc;
                hipLaunchKernelGGL(cl_add_acc, dim3(rb), dim3(256), 0, s, n_rep, row_acc, acc);
            } else {
                cl_cluster_row(a, code, count, sweeps, s);
                int rc = cm_ising_measure(spins, ndim, dims, n_rep, e_run, m_run, s);
                if (rc != CM_OK) return rc;
            }
            if (eager)
                for (int r = 0; r < n_rep; ++r) sweeps_host[r] += (uint64_t)count;
        }
        hipLaunchKernelGGL(cl_record, dim3(rb), dim3(256), 0, s, n_rep, n_meas,
                           (const int64_t*)meas_d, e_run, m_run, e_out, m_out, acc, a.wsize,
                           a.stats, any_sw ? 1 : 0, aux_out);
        hipLaunchKernelGGL(cl_next, dim3(1), dim3(64), 0, s, meas_d);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    };
    bool graphed;
    const bool capturable = !any_local || ising_rows_capturable(ndim, dims, n_rep);
    if (int rc = repeat_measurements(n_meas, capturable, st, &graphed, measurement)) return rc;
    if (graphed)
        for (int r = 0; r < n_rep; ++r) sweeps_host[r] += (uint64_t)(per_meas * n_meas);
    return CM_OK;
}

int64_t cm_potts_cluster_workspace_bytes(int32_t ndim, const int32_t* dims, int32_t n_rep,
                                         int32_t q)
{
    int64_t N = 1;
    for (int a = 0; a < ndim; ++a) N *= dims[a];
    return cl_layout(N, n_rep, cm_potts_workspace_bytes(n_rep), q).total;
}

int cm_potts_cluster_run(void* spins, int32_t ndim, const int32_t* dims, int32_t n_rep, int32_t q,
                         const double* betas, const uint32_t* keys, const uint32_t* stream_ids,
                         uint64_t* sweeps, const int64_t* schedule, int32_t n_rows, int64_t n_meas,
                         int32_t* e_run, int32_t* counts_run, int32_t* e_out, int32_t* counts_out,
                         int64_t* aux_out, void* workspace, int64_t workspace_bytes, void* stream)
{
    int64_t N;
    bool any_cluster, any_sw;
    if (int rc = cl_check(ndim, dims, n_rep, n_meas, schedule, n_rows, &N, &any_cluster, &any_sw))
        return rc;
    CM_REQUIRE(q >= 2 && q <= 127, "q must be in [2, 127], got %d", q);
    const ClLayout lay = cl_layout(N, n_rep, cm_potts_workspace_bytes(n_rep), q);
    CM_REQUIRE(workspace_bytes >= lay.total, "workspace too small");
    if (n_meas == 0) return CM_OK;
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    std::vector<uint64_t> tadd(n_rep);
    for (int r = 0; r < n_rep; ++r) {
        CM_REQUIRE(std::isfinite(betas[r]), "beta must be finite");
        tadd[r] = cl_threshold(1.0 - std::exp(-betas[r]));
    }
    CM_HIP_CHECK(upload_async(ws + lay.tadd, tadd.data(), tadd.size() * 8, st));
    ClArgs a{};
    cl_args(a, ws, lay, spins, ndim, dims, N, n_rep, q, keys, stream_ids, sweeps);
    int64_t* acc = (int64_t*)(ws + lay.acc);
    int32_t* row_acc = (int32_t*)(ws + lay.rowacc);
    int32_t* row_e = (int32_t*)(ws + lay.rowe);
    int32_t* row_c = (int32_t*)(ws + lay.rowm);
    const int rb = (n_rep + 255) / 256;
    int64_t* meas_d = (int64_t*)(ws + lay.meas);
    hipError_t lds_err;
    if (cl_lds_try(a, schedule, n_rows, n_meas, sweeps, e_run, counts_run, e_out, counts_out,
                   aux_out, st, &lds_err)) {
        CM_HIP_CHECK(lds_err);
        return CM_OK;
    }
    hipLaunchKernelGGL(cl_zero64, dim3((2 * n_rep + 255) / 256), dim3(256), 0, st, a.stats, 2 * n_rep);
    hipLaunchKernelGGL(cl_set64, dim3(1), dim3(64), 0, st, meas_d, (int64_t)0);
    // the local rows' tables are uploaded once per call
    for (int row = 0; row < n_rows; ++row) {
        if (schedule[2 * row] != CODE_METROPOLIS && schedule[2 * row] != CODE_HEATBATH) continue;
        if (int rc = potts_prepare(ndim, dims, n_rep, betas, ws + lay.tab, lay.parent - lay.tab,
                                   stream))
            return rc;
        break;
    }
    auto measurement = [&](hipStream_t s, bool) -> int {
        hipLaunchKernelGGL(cl_zero64, dim3(rb), dim3(256), 0, s, acc, n_rep);
        hipLaunchKernelGGL(cl_zero64, dim3(rb), dim3(256), 0, s, a.wsize, n_rep);
        for (int row = 0; row < n_rows; ++row) {
            const int64_t code = schedule[2 * row], count = schedule[2 * row + 1];
            if (code == CODE_METROPOLIS || code == CODE_HEATBATH) {
                const int64_t one[2] = {code, count};
                int rc = potts_rows(spins, ndim, dims, n_rep, q, keys, stream_ids, sweeps, one, 1,
                                    1, e_run, counts_run, row_e, row_acc, row_c, ws + lay.tab,
                                    lay.parent - lay.tab, 0, 0, nullptr, s);
                if (rc != CM_OK) return rc;
                hipLaunchKernelGGL(cl_add_acc, dim3(rb), dim3(256), 0, s, n_rep, row_acc, acc);
            } else {
                cl_cluster_row(a, code, count, sweeps, s);
                int rc = cm_potts_measure(spins, ndim, dims, n_rep, q, e_run, counts_run, s);
                if (rc != CM_OK) return rc;
            }
        }
        hipLaunchKernelGGL(cl_record_potts, dim3(rb), dim3(256), 0, s, n_rep, q, n_meas,
                           (const int64_t*)meas_d, e_run, counts_run, e_out, counts_out, acc,
                           a.wsize, a.stats, any_sw ? 1 : 0, aux_out);
        hipLaunchKernelGGL(cl_next, dim3(1), dim3(64), 0, s, meas_d);
        CM_HIP_CHECK(hipGetLastError());
        return CM_OK;
    };
    bool graphed;
    return repeat_measurements(n_meas, true, st, &graphed, measurement);
}

int64_t cm_table_workspace_bytes(int64_t N, int32_t n_rep, int32_t q)
{
    return cl_layout(N, n_rep, (int64_t)n_rep * SEQ_TAB * 8, q > 0 ? q : 1).total;
}

int cm_ising_table_run(void* spins, int64_t N, int32_t z, const int32_t* nbr, int32_t n_rep,
                       const double* betas, double h, const uint32_t* keys,
                       const uint32_t* stream_ids, uint64_t* sweeps, const int64_t* schedule,
                       int32_t n_rows, int64_t n_meas, int32_t* e_run, int32_t* m_run,
                       int32_t* e_out, int32_t* m_out, int64_t* aux_out, void* workspace,
                       int64_t workspace_bytes, void* stream)
{
    return table_run(0, spins, N, z, nbr, n_rep, betas, h, keys, stream_ids, sweeps, schedule,
                     n_rows, n_meas, e_run, m_run, e_out, m_out, aux_out, workspace,
                     workspace_bytes, (hipStream_t)stream);
}

int cm_potts_table_run(void* spins, int64_t N, int32_t z, const int32_t* nbr, int32_t n_rep,
                       int32_t q, const double* betas, const uint32_t* keys,
                       const uint32_t* stream_ids, uint64_t* sweeps, const int64_t* schedule,
                       int32_t n_rows, int64_t n_meas, int32_t* e_run, int32_t* counts_run,
                       int32_t* e_out, int32_t* counts_out, int64_t* aux_out, void* workspace,
                       int64_t workspace_bytes, void* stream)
{
    CM_REQUIRE(q >= 2 && q <= 127, "q must be in [2, 127], got %d", q);
    return table_run(q, spins, N, z, nbr, n_rep, betas, 0.0, keys, stream_ids, sweeps, schedule,
                     n_rows, n_meas, e_run, counts_run, e_out, counts_out, aux_out, workspace,
                     workspace_bytes, (hipStream_t)stream);
}

int cm_table_measure(const void* spins, int64_t N, int32_t z, const int32_t* nbr, int32_t n_rep,
                     int32_t q, int32_t* e_run, int32_t* mc_run, void* stream)
{
    CM_REQUIRE(N >= 2 && z >= 2 && z <= SEQ_ZMAX && z % 2 == 0 && n_rep >= 1 && n_rep <= 65535,
               "bad lattice or replica count");
    CM_REQUIRE(q == 0 || (q >= 2 && q <= 127), "q must be 0 (Ising) or in [2, 127]");
    SeqArgs sa{};
    sa.spins = (int8_t*)spins;
    sa.nbr = nbr;
    sa.N = (uint32_t)N;
    sa.z = z;
    sa.R = n_rep;
    sa.q = q;
    sa.e_run = e_run;
    sa.m_run = q ? nullptr : mc_run;
    sa.counts = q ? mc_run : nullptr;
    hipLaunchKernelGGL(seq_measure, dim3(n_rep), dim3(256), 0, (hipStream_t)stream, sa);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

int cm_set_option(const char* name, int64_t value, int64_t* previous)
{
    CM_REQUIRE(name != nullptr, "option name missing");
    int64_t* slot = nullptr;
    if (std::strcmp(name, "graph_min_meas") == 0) slot = &g_graph_min_meas;
    else if (std::strcmp(name, "cluster_lds_max_sites") == 0) slot = &g_cluster_lds_max_sites;
    else if (std::strcmp(name, "ea_fused") == 0) slot = &g_ea_fused;
    else if (std::strcmp(name, "ea_order") == 0) slot = &g_ea_order;
    CM_REQUIRE(slot != nullptr, "unknown option %s", name);
    if (previous) *previous = *slot;
    *slot = value;
    return CM_OK;
}

}  // extern "C"
