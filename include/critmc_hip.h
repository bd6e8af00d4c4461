/*
 * critmc_hip.h -- C-ABI of libcritmc_hip.so, the MI355X (gfx950) lattice-update
 * engine behind critmc's kernel seam.
 *
 * The reference has no FFI: its seam is a Python call into numba,
 * Model.advance -> kernels/<model>.run (critmc/models.py:58-68).  Each entry
 * point below replaces one of those compiled calls and keeps its conventions:
 * the caller owns every buffer (state, outputs, scratch), state and running
 * totals are updated in place, and argument checks report errors instead of
 * raising (return 0 = ok, -1 = invalid argument, -2 = unsupported, -3 = HIP
 * error; cm_last_error() gives the text).
 *
 * Pointers named *_dev are device memory; everything else is host memory.
 * `stream` is a hipStream_t (NULL = default stream); kernels are launched on
 * it asynchronously unless stated otherwise.
 *
 * Random numbers: Philox4x32-10, key (k0, k1) per chain, counter
 * (block, sweep_lo, sweep_hi16 | colour << 16 | tag << 24, stream_id), where
 * `sweep` is the chain's own count of local sweeps -- the replacement of the
 * numpy Generator argument (critmc/rng.py:49-66), reproducible under any
 * chunking (critmc tests/test_sim_chunks.py).
 */
#ifndef CRITMC_HIP_H
#define CRITMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------ */

/* ABI version of this header. */
int32_t cm_version(void);

/* Text of the last error raised on the calling thread. */
const char *cm_last_error(void);

/* Copy n device-to-device spans in one launch: spans_dev holds n records
 * (dst, src, bytes) of uint64, with dst and src 4-byte aligned and bytes a
 * multiple of 4, and the spans must not overlap.  No counterpart in the
 * reference (whose runner copies one chain's numpy arrays at a time); the
 * serving loop hands every job of a pool its slice of the time series with it
 * instead of one copy per job and series. */
int cm_copy_spans(const void *spans_dev, int32_t n, void *stream);

/* Device properties used by the host-side tuner. */
int cm_device_info(int32_t *cu_count, int64_t *lds_per_block, int32_t *arch_major,
                   int32_t *arch_minor);

/* Raw Philox4x32-10 blocks with counters (x0 + b, y, z, w), b < n_blocks,
 * written as 4 u32 per block; replaces the numpy bit generator draws
 * (critmc/rng.py:66) for testing the stream contract. */
int cm_philox_fill(uint32_t *out_dev, int64_t n_blocks, uint32_t k0, uint32_t k1, uint32_t x0,
                   uint32_t y, uint32_t z, uint32_t w, void *stream);

/* The same function on the host (ctr[4], key[2] -> out[4]). */
void cm_philox_host(const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* ---- Ising (critmc/kernels/ising.py) ------------------------------------- */

/* Bytes of caller-owned device scratch cm_ising_run needs for n_rep chains. */
int64_t cm_ising_workspace_bytes(int32_t n_rep);

/* The acceptance-class table cm_ising_run builds for one chain (64 u32):
 * critmc/kernels/ising.py:47-65 (metropolis_table, heatbath_table) as
 * integer thresholds T = ceil(p 2^32).  For tests and inspection. */
int cm_ising_tables(double beta, double h, int32_t z, uint32_t *tab_out);

/* Initial int8 spins [n_rep][N]: start 0 = random (Philox tag INIT),
 * 1 = all +1, 2 = all -1.  Replaces Ising.initial_state (critmc/models.py:113-121)
 * for device-side batches. */
int cm_ising_init(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep,
                  const uint32_t *keys_dev, const uint32_t *stream_ids_dev, int32_t start,
                  void *stream);

/* Exact E_b = -sum s_i s_j and M = sum s_i per replica (int32):
 * critmc/kernels/ising.py:31-44 (bond_energy, magnetization). */
int cm_ising_measure(const void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep,
                     int32_t *e_run_dev, int32_t *m_run_dev, void *stream);

/* int8 [n_rep][N] <-> bit-packed [n_rep][2][rows][ceil(L0/64)] u32;
 * direction 0 packs, 1 unpacks.  No reference counterpart (layout conversion). */
int cm_ising_convert(void *spins_i8_dev, void *spins_packed_dev, int32_t ndim,
                     const int32_t *dims, int32_t n_rep, int32_t direction, void *stream);

/* Apply `schedule` (n_rows rows of (update code, count); codes 0 = metropolis,
 * 2 = heatbath) n_meas times to n_rep chains of a periodic hypercubic lattice
 * with even extents, recording E_b, M and the accepted count after every
 * application into e_out/m_out/acc_out [n_rep][n_meas] (int32).
 * Replaces critmc/kernels/ising.py:173-237 (run) for checkerboard sweeps.
 *   layout  0 = int8 spins, 1 = packed (see cm_ising_convert)
 *   betas   host [n_rep]; h is the uniform field
 *   keys_dev [n_rep][2], stream_ids_dev [n_rep], sweeps_dev [n_rep] (in/out);
 *   sweeps_host [n_rep] mirrors sweeps_dev (read by the plain form)
 *   e_run_dev / m_run_dev [n_rep]: running E_b, M (in/out)
 *   workspace_dev: >= cm_ising_workspace_bytes(n_rep) bytes
 *   form    0 = auto, 1 = plain (one chain, one launch per half sweep),
 *           2 = batched (one launch, LDS-resident bit-packed replicas; 2D with
 *           L0 <= 64 or L0 % 64 == 0), 3 = wide (one launch per half sweep
 *           over all replicas, one lane per 32-site word: 2D with L0 % 64 == 0
 *           in either layout, any other lattice with int8 spins through the
 *           plain form's words; auto picks it for few replicas of >= 65536
 *           sites, where one workgroup per replica would leave CUs idle, and
 *           for every int8 lattice the batched form does not take)
 *   block, grid: launch shape of the batched form (0 = default)
 *   n_meas_rep_dev: null, or int64 [n_rep] caps (batched and wide forms): chain r performs
 *           min(n_meas, n_meas_rep[r]) measurements (continuous batching of
 *           jobs with different remaining work)
 *   prepared: null (the tables are built from betas and h and uploaded on
 *           `stream` by this call, as the reference builds them in every run,
 *           kernels/ising.py:200), or the info[2] record cm_ising_prepare
 *           wrote for these betas, h and update codes into this workspace
 *           (no host work per call; betas is then not read)
 * The plain form synchronises `stream` before returning. */
int cm_ising_run(void *spins_dev, int32_t layout, int32_t ndim, const int32_t *dims,
                 int32_t n_rep, const double *betas, double h, const uint32_t *keys_dev,
                 const uint32_t *stream_ids_dev, uint64_t *sweeps_dev, const uint64_t *sweeps_host,
                 const int64_t *schedule, int32_t n_rows, int64_t n_meas, int32_t *e_run_dev,
                 int32_t *m_run_dev, int32_t *e_out_dev, int32_t *m_out_dev, int32_t *acc_out_dev,
                 void *workspace_dev, int64_t workspace_bytes, int32_t form, int32_t block,
                 int32_t grid, const int64_t *n_meas_rep_dev, const int32_t *prepared,
                 void *stream);

/* Build the class tables of n_rep chains (critmc/kernels/ising.py:47-65,
 * metropolis_table / heatbath_table, which the reference rebuilds at the top
 * of every run, :200) and upload them on `stream` into workspace_dev, once for
 * any number of later cm_ising_run calls with the same betas, h and update
 * codes.  Writes info[0] = the largest acceptance-class count a row of this
 * schedule reads, info[1] = 1 if every row is Metropolis and every table is the
 * zero-field z = 4 one; pass info as cm_ising_run's `prepared`.  The workspace
 * must not be written by anything else in between. */
int cm_ising_prepare(int32_t ndim, const int32_t *dims, int32_t n_rep, const double *betas,
                     double h, const int64_t *schedule, int32_t n_rows, void *workspace_dev,
                     int64_t workspace_bytes, int32_t *info, void *stream);

/* Bytes of caller-owned device scratch cm_ising_cluster_run needs (tables plus
 * parent, size, queue and mark arrays of n_rep x N int32). */
int64_t cm_ising_cluster_workspace_bytes(int32_t ndim, const int32_t *dims, int32_t n_rep);

/* Any Ising schedule, including cluster rows: code 3 = Wolff (one cluster per
 * repetition), 4 = Swendsen-Wang (one update per repetition), mixed with
 * Metropolis / heat-bath rows; replaces critmc/kernels/ising.py:173-237 (run)
 * together with wolff_update / sw_update (:110-170).  Int8 spins [n_rep][N].
 * Records E_b, M (int32 [n_rep][n_meas]) and the reference's aux columns
 * aux_out int64 [n_rep][n_meas][4] = (accepted, Wolff cluster sizes summed over
 * the measurement, sum |C|^2 and cluster count of the measurement's last
 * Swendsen-Wang update).  Every repetition of any row advances the chain's
 * sweep counter by one; sweeps_host mirrors sweeps_dev and is advanced here.
 * Cluster rows need h == 0.  Local rows use the same kernels as cm_ising_run. */
int cm_ising_cluster_run(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep,
                         const double *betas, double h, const uint32_t *keys_dev,
                         const uint32_t *stream_ids_dev, uint64_t *sweeps_dev,
                         uint64_t *sweeps_host, const int64_t *schedule, int32_t n_rows,
                         int64_t n_meas, int32_t *e_run_dev, int32_t *m_run_dev,
                         int32_t *e_out_dev, int32_t *m_out_dev, int64_t *aux_out_dev,
                         void *workspace_dev, int64_t workspace_bytes, void *stream);

/* ---- Potts (critmc/kernels/potts.py) -------------------------------------- */

/* Bytes of caller-owned device scratch cm_potts_run needs for n_rep chains. */
int64_t cm_potts_workspace_bytes(int32_t n_rep);

/* Initial int8 states [n_rep][N] in 0..q-1: start 0 = random (Philox tag
 * INIT, floor(x q / 2^32)), 1 = all 0.  Replaces Potts.initial_state
 * (critmc/models.py:177-183) for device-side batches. */
int cm_potts_init(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t q,
                  const uint32_t *keys_dev, const uint32_t *stream_ids_dev, int32_t start,
                  void *stream);

/* E = -#satisfied bonds and counts[q] per replica: critmc/kernels/potts.py:27-41. */
int cm_potts_measure(const void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep,
                     int32_t q, int32_t *e_run_dev, int32_t *counts_run_dev, void *stream);

/* Apply `schedule` (codes 0 = metropolis, 2 = heatbath) n_meas times to n_rep
 * q-state Potts chains (2 <= q <= 127), recording E, accepted [n_rep][n_meas]
 * and counts [n_rep][n_meas][q] after every application.
 * Replaces critmc/kernels/potts.py:170-235 (run) for checkerboard sweeps.
 * e_run_dev [n_rep] and counts_run_dev [n_rep][q] are the running totals (in/out);
 * grid 0 picks the form, grid > 0 runs one workgroup per replica (states in
 * LDS up to 96 KB), grid < 0 the wide form (one launch per half sweep over all
 * replicas); every form gives the same chains.  betas null: the tables are
 * already in the workspace (cm_potts_prepare) and no host work is done.
 * Other arguments as cm_ising_run (batched form). */
int cm_potts_run(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t q,
                 const double *betas, const uint32_t *keys_dev, const uint32_t *stream_ids_dev,
                 uint64_t *sweeps_dev, const int64_t *schedule, int32_t n_rows, int64_t n_meas,
                 int32_t *e_run_dev, int32_t *counts_run_dev, int32_t *e_out_dev,
                 int32_t *acc_out_dev, int32_t *counts_out_dev, void *workspace_dev,
                 int64_t workspace_bytes, int32_t block, int32_t grid,
                 const int64_t *n_meas_rep_dev, void *stream);

/* Build the tables of n_rep Potts chains (Metropolis thresholds and heat-bath
 * weights, critmc/kernels/potts.py:194-200, which the reference rebuilds at
 * the top of every run) and upload them on `stream` into workspace_dev, once
 * for any number of later cm_potts_run calls with betas = null on the same
 * lattice dimension.  The workspace must not be written by anything else in
 * between. */
int cm_potts_prepare(int32_t ndim, const int32_t *dims, int32_t n_rep, const double *betas,
                     void *workspace_dev, int64_t workspace_bytes, void *stream);

/* Bytes of device scratch cm_potts_cluster_run needs. */
int64_t cm_potts_cluster_workspace_bytes(int32_t ndim, const int32_t *dims, int32_t n_rep,
                                         int32_t q);

/* Any Potts schedule, including Wolff (3) and Swendsen-Wang (4) rows:
 * critmc/kernels/potts.py:170-235 with wolff_update / sw_update (:106-167).
 * Records E and counts like cm_potts_run (int32) and the reference's aux
 * columns aux_out int64 [n_rep][n_meas][4] (accepted, Wolff sizes, sum |C|^2
 * and cluster count of the last Swendsen-Wang update).  Every repetition of
 * any row advances the chain's sweep counter by one. */
int cm_potts_cluster_run(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep,
                         int32_t q, const double *betas, const uint32_t *keys_dev,
                         const uint32_t *stream_ids_dev, uint64_t *sweeps_dev,
                         const int64_t *schedule, int32_t n_rows, int64_t n_meas,
                         int32_t *e_run_dev, int32_t *counts_run_dev, int32_t *e_out_dev,
                         int32_t *counts_out_dev, int64_t *aux_out_dev, void *workspace_dev,
                         int64_t workspace_bytes, void *stream);

/* ---- Edwards-Anderson parallel tempering (critmc/kernels/glass.py) -------- */

/* A batch holds n_samples disorder samples of the same dimension (2 or 3) and
 * possibly different even extents L[s]; sample s owns int8 spins
 * [2 sets][m configurations][L^d] at spin_off[s] and forward-bond couplings
 * [d][L^d] at j_off[s] (bond a*N + i joins i and its +e_a neighbour,
 * critmc/lattice.py:136-138).  jtype 0: int8 +-1 couplings (disorder
 * "bimodal"), int32 energies and energy series; jtype 1: FP64 couplings (e.g.
 * disorder "gaussian", critmc/glass.py:60-64), FP64 energies and series. */

/* Bytes of caller-owned device scratch cm_ea_pt_run needs; total_sites is
 * sum_s L[s]^d. */
int64_t cm_ea_workspace_bytes(int32_t n_samples, int32_t m, int64_t total_sites);

/* e^x as the tempering swap test evaluates it (fixed FP64 operation sequence). */
double cm_det_exp(double x);

/* 1 if cm_ea_pt_run runs a batch with these parameters multi-spin coded
 * (+-1 couplings, m <= 32, largest extent max_L <= 32 in 2D or <= 16 in 3D),
 * which fixes its uniform mapping; 0 otherwise.  No counterpart in the
 * reference, whose kernels have one mapping. */
int cm_ea_multispin(int32_t jtype, int32_t m, int32_t ndim, int32_t max_L);

/* Random +-1 start of every configuration (Philox tag INIT, stream r*m + c,
 * key = the sample's).  Replaces the spin draw of ParallelTempering.__init__
 * (critmc/glass.py:148) for device-side batches. */
int cm_ea_init(void *spins_dev, const int32_t *L_dev, const int64_t *spin_off_dev,
               int32_t n_samples, int32_t m, int32_t ndim, const uint32_t *keys_dev,
               void *stream);

/* E = -sum_b J_b s_i s_j of every configuration into energies_dev [S][2][m]:
 * critmc/kernels/glass.py:33-38 (energy). */
int cm_ea_energy(void *spins_dev, const void *J_dev, int32_t jtype, const int32_t *L_dev,
                 const int64_t *spin_off_dev, const int64_t *j_off_dev, int32_t n_samples,
                 int32_t m, int32_t ndim, void *energies_dev, void *stream);

/* n_steps parallel-tempering steps of every sample: sweeps_per_step sweeps
 * (code 0 = Metropolis as the reference, 2 = heat bath) of every
 * configuration at its temperature, neighbour swaps k = 0..m-2 for each set,
 * round-trip bookkeeping, then per temperature q, q_l, |q(k_min)|^2 into the
 * log-binned accumulators acc [S][n_bins][m][7] (bin of the call's step t:
 * floor(log2(t+1)); columns e, q_l, |q|, q^2, q^4, |q(k)|^2, count) and, when
 * the pointers are non-null, the series e_series [S][2][m][n_steps] (energy
 * at temperature k; int32 or FP64 per jtype) and q_series [S][m][n_steps] (N q).
 * With FP64 couplings the energies are recomputed from the spins at the start
 * of every bin after the first (critmc/kernels/glass.py:116-122).
 * Replaces critmc/kernels/glass.py:80-163 (pt_run); state arrays are
 * [S][2][m] (perm: temperature -> configuration; kofc its inverse), trips
 * [S][2], label_hist [S][2][m][2], swap_acc / swap_try [S][2][m-1].
 * betas_host / betas_dev: [S][m], hot -> cold.  step0: the global PT step of
 * the first step (keys the swap draws).  L_host mirrors L_dev.
 * Uniforms: with +-1 couplings, m <= 32 and L <= 32 (2D) or L <= 16 (3D) the
 * sweep is multi-spin coded (one bit per configuration): colour site k of set r draws
 * its 32 bit planes from Philox blocks 8k .. 8k+7 (tag EA, stream r, the
 * set's sweep counter = that of its configuration 0), and configuration c
 * takes bit c of every plane, most significant first; otherwise configuration
 * c of set r uses word k&3 of block k>>2 (stream r*m + c, its own counter). */
int cm_ea_pt_run(void *spins_dev, const void *J_dev, int32_t jtype, const int32_t *L_dev,
                 const int64_t *spin_off_dev, const int64_t *j_off_dev, int32_t n_samples,
                 int32_t m, int32_t ndim, const int32_t *L_host, const double *betas_host,
                 const double *betas_dev, const uint32_t *keys_dev, uint64_t *sweeps_dev,
                 void *energies_dev, int32_t *perm_dev, int32_t *kofc_dev, int32_t *labels_dev,
                 int32_t *trips_dev, int32_t *label_hist_dev, int32_t *swap_acc_dev,
                 int32_t *swap_try_dev, double *acc_dev, int32_t n_bins, void *e_series_dev,
                 int32_t *q_series_dev, int32_t code, int32_t sweeps_per_step, uint64_t step0,
                 int64_t n_steps, void *workspace_dev, int64_t workspace_bytes, int32_t block,
                 void *stream);

/* ---- O(n) vector model, n = 2 (XY) and 3 (critmc/kernels/on.py) ---------- */

/* Bytes of caller-owned device scratch cm_on_run needs: FP64 partial sums,
 * and for XY on square lattices a second spin buffer of n_rep * N * 2 floats
 * (the fused over-relaxation launch writes out of place). */
int64_t cm_on_workspace_bytes(int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t n);

/* O(n) spins live on the device in the colour-split layout
 * [n_rep][2 colours][N/2][n] FP32: site i (x fastest) of colour
 * c = (x0 + x1 + x2) mod 2 is entry i >> 1 of plane c. */

/* Site order [n_rep][N][n] <-> colour-split layout, out of place;
 * direction 0: src in site order, dst split; 1: the reverse. */
int cm_on_convert(const void *src_dev, void *dst_dev, int32_t ndim, const int32_t *dims,
                  int32_t n_rep, int32_t n, int32_t direction, void *stream);

/* Random FP32 unit vectors (XY: angle uniform in [0, 2 pi); n = 3:
 * isotropic), Philox tag INIT, drawn per site index and stored in the
 * colour-split layout.  Replaces ON.initial_state (critmc/models.py:248-257)
 * for device-side batches. */
int cm_on_init(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t n,
               const uint32_t *keys_dev, const uint32_t *stream_ids_dev, void *stream);

/* Bytes of the cluster workspace cm_on_run needs for Wolff / Swendsen-Wang rows. */
int64_t cm_on_cluster_workspace_bytes(int32_t ndim, const int32_t *dims, int32_t n_rep);

/* Apply `schedule` (codes 0 = metropolis, 2 = heatbath, 3 = wolff,
 * 4 = swendsen-wang, 5 = overrelax) n_meas times to n_rep FP32 O(n) chains;
 * after every application renormalise in FP64 and record FP64 E
 * [n_rep][n_meas], M [n_rep][n_meas][n], stiffness sums c_a / s_a
 * [n_rep][n_meas][ndim] and int32 accepted Metropolis moves.
 * betas_dev / steps_dev: FP32 [n_rep] (inverse temperature, Metropolis step).
 * n_meas_rep_dev: null or int64 [n_rep] per-chain caps as in cm_ising_run
 * (local updates only).  aux_out_dev: null or int64 [n_rep][n_meas][4], the
 * reference's aux columns (accepted, Wolff sizes, sum |C|^2 and cluster count
 * of the last Swendsen-Wang update).  cluster_ws_dev: null, or at least
 * cm_on_cluster_workspace_bytes for schedules with cluster rows (the embedded
 * Ising updates of critmc/kernels/on.py:264-345).
 * A call of local updates only, with aux_out_dev null, on a lattice whose
 * spins fit in 158 KB (N * n * 4 bytes; XY up to 128^2) runs as one launch,
 * one workgroup per replica with its spins in LDS; other calls launch per half
 * sweep.  Both
 * forms give the same spins and series.
 * Replaces critmc/kernels/on.py:348-418 (run). */
int cm_on_run(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t n,
              const float *betas_dev, const float *steps_dev, const uint32_t *keys_dev,
              const uint32_t *stream_ids_dev, uint64_t *sweeps_dev, const int64_t *schedule,
              int32_t n_rows, int64_t n_meas, int32_t *acc_out_dev, double *e_out_dev,
              double *m_out_dev, double *sc_out_dev, double *ss_out_dev, void *workspace_dev,
              int64_t workspace_bytes, const int64_t *n_meas_rep_dev, int64_t *aux_out_dev,
              void *cluster_ws_dev, int64_t cluster_ws_bytes, void *stream);

/* ---- Flat-histogram walkers (critmc/kernels/flathist.py) ------------------ */

/* n_moves Wang-Landau moves of each of n_rep independent walkers
 * (critmc/kernels/flathist.py:22-74, wl_ising / wl_potts), one lane per walker.
 * q = 0: Ising, levels (E_b + B)/4; q >= 2: Potts, levels E + B; n_levels must
 * be B/2 + 1 or B + 1.  Per walker: int8 spins [N], idx (level, in/out),
 * ln_g and int64 hist [n_levels] (in/out), window [lo, hi], Philox key and
 * stream, moves (the walker's move count, in/out), ln_f, mode_1t (nonzero:
 * ln f = n_levels_1t / (t + 1) per move t, Belardinelli-Pereyra); active
 * (null = all) skips walkers whose entry is 0.  Acceptance uses the
 * fixed-operation e^x (cm_det_exp). */
int cm_wl_run(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t q,
              int32_t *idx_dev, double *ln_g_dev, int64_t *hist_dev, int32_t n_levels,
              const int32_t *lo_dev, const int32_t *hi_dev, const uint32_t *keys_dev,
              const uint32_t *stream_ids_dev, uint64_t *moves_dev, const double *ln_f_dev,
              const int32_t *mode_1t_dev, const double *n_levels_1t_dev, const int32_t *active_dev,
              int64_t n_moves, void *stream);

/* Multicanonical production with fixed weights ln_w [n_rep][n_levels]
 * (critmc/kernels/flathist.py:77-132, muca_ising / muca_potts): n_meas
 * measurements of moves_per_meas moves per walker, recording the level
 * (int32 idx_out [n_rep][n_meas]) and M (Ising, order = M [n_rep]) or
 * sum_k n_k^2 (Potts, order = counts [n_rep][q]) into int64 ord_out. */
int cm_muca_run(void *spins_dev, int32_t ndim, const int32_t *dims, int32_t n_rep, int32_t q,
                int32_t *idx_dev, int32_t *order_dev, const double *ln_w_dev, int64_t *hist_dev,
                int32_t n_levels, const int32_t *lo_dev, const int32_t *hi_dev,
                const uint32_t *keys_dev, const uint32_t *stream_ids_dev, uint64_t *moves_dev,
                int64_t n_meas, int64_t moves_per_meas, int32_t *idx_out_dev, int64_t *ord_out_dev,
                void *stream);

/* Library tuning options (no reference counterpart; results never depend on them).
 *   "graph_min_meas": schedules with cluster rows that run measurement by
 *       measurement capture one measurement's launches into a hipGraph and
 *       replay it when a call has at least this many measurements (<= 0, the
 *       default: never; measured no faster on MI355X).
 *   "cluster_lds_max_sites": schedules made only of Wolff / Swendsen-Wang rows
 *       on lattices of at most this many sites (and at most 8192) run every
 *       update of the call in one launch, each replica in one workgroup's LDS
 *       (default 8192; 0: always the multi-launch kernels).
 *   "ea_fused": multi-spin EA tempering (+-J, m <= 32) runs every PT step of
 *       a call in one launch, one workgroup per sample (1, the default), or
 *       three launches per step (0).
 *   "ea_order": dispatch order of the fused launch's samples: largest first
 *       (0, the default) or the larger half largest first and the smaller
 *       half smallest first (1).
 * Sets `name` to value; *previous (if not null) receives the old value. */
int cm_set_option(const char *name, int64_t value, int64_t *previous);

/* ---- Chains on any lattice (neighbour table) ----------------------------- */

/* Bytes of device scratch for n_rep chains of n_sites sites (q = 0: Ising,
 * else the number of Potts states). */
int64_t cm_table_workspace_bytes(int64_t n_sites, int32_t n_rep, int32_t q);

/* Any Ising schedule on a lattice given by its neighbour table nbr_dev
 * int32 [n_sites][z] in pair form (column 2a = +e_a, 2a+1 = -e_a; z even, <= 8):
 * odd extents, the triangular lattice (critmc/lattice.py:173-213), or any
 * lattice with `metropolis-random` rows.  Replaces critmc/kernels/ising.py:
 * 173-237 (run) with the reference's own site orders: metropolis (0) and
 * heatbath (2) visit sites 0..N-1 in turn, metropolis-random (1) makes N
 * attempts at random sites (:68-107), one lane per chain; wolff (3) and sw (4)
 * (:110-170) take neighbours and bonds (a N + i joins i and nbr[i][2a]) from
 * the table.  Records E_b, M (int32 [n_rep][n_meas]) and aux_out int64
 * [n_rep][n_meas][4] as cm_ising_cluster_run; e_run_dev / m_run_dev and
 * sweeps_dev (one value per repetition of any row) advance.  Cluster rows need
 * h == 0. */
int cm_ising_table_run(void *spins_dev, int64_t n_sites, int32_t z, const int32_t *nbr_dev,
                       int32_t n_rep, const double *betas, double h, const uint32_t *keys_dev,
                       const uint32_t *stream_ids_dev, uint64_t *sweeps_dev,
                       const int64_t *schedule, int32_t n_rows, int64_t n_meas,
                       int32_t *e_run_dev, int32_t *m_run_dev, int32_t *e_out_dev,
                       int32_t *m_out_dev, int64_t *aux_out_dev, void *workspace_dev,
                       int64_t workspace_bytes, void *stream);

/* The q-state Potts counterpart (critmc/kernels/potts.py:44-235): records E,
 * counts (int32 [n_rep][n_meas][q]) and aux_out; counts_run_dev [n_rep][q]. */
int cm_potts_table_run(void *spins_dev, int64_t n_sites, int32_t z, const int32_t *nbr_dev,
                       int32_t n_rep, int32_t q, const double *betas, const uint32_t *keys_dev,
                       const uint32_t *stream_ids_dev, uint64_t *sweeps_dev,
                       const int64_t *schedule, int32_t n_rows, int64_t n_meas,
                       int32_t *e_run_dev, int32_t *counts_run_dev, int32_t *e_out_dev,
                       int32_t *counts_out_dev, int64_t *aux_out_dev, void *workspace_dev,
                       int64_t workspace_bytes, void *stream);

/* Exact energy and magnetisation (q = 0: Ising E_b, M into mc_run_dev [n_rep])
 * or occupation counts (q >= 2: mc_run_dev [n_rep][q]) from the neighbour table:
 * critmc/kernels/ising.py:31-44, kernels/potts.py:27-41. */
int cm_table_measure(const void *spins_dev, int64_t n_sites, int32_t z, const int32_t *nbr_dev,
                     int32_t n_rep, int32_t q, int32_t *e_run_dev, int32_t *mc_run_dev,
                     void *stream);

/* O(n) chains (n = 2, 3) on a neighbour-table lattice, one lane per chain, in
 * FP64: replaces k_on.run (critmc/kernels/on.py:348-418) on lattices without a
 * checkerboard and for metropolis-random (:112-137), with heat bath
 * (:186-261), over-relaxation (:140-183) and Wolff / Swendsen-Wang rows
 * (:264-345).  spins_dev: FP64 [n_rep][n_sites][n] in site order; betas,
 * steps: host [n_rep].  Per measurement every spin is renormalised and
 * e_out_dev [n_rep][n_meas], m_out_dev [..][n], c_out_dev / s_out_dev
 * [..][z/2] (stiffness sums per bond axis) and aux_out_dev [..][4] (accepted,
 * Wolff sizes, sum |C|^2, clusters) are written.  Uniforms: see seq_on.hip. */
int64_t cm_on_table_workspace_bytes(int64_t n_sites, int32_t n_rep);
int cm_on_table_run(void *spins_dev, int64_t n_sites, int32_t z, const int32_t *nbr_dev,
                    int32_t n_rep, int32_t n, const double *betas, const double *steps,
                    const uint32_t *keys_dev, const uint32_t *stream_ids_dev,
                    uint64_t *sweeps_dev, const int64_t *schedule, int32_t n_rows,
                    int64_t n_meas, double *e_out_dev, double *m_out_dev, double *c_out_dev,
                    double *s_out_dev, int64_t *aux_out_dev, void *workspace_dev,
                    int64_t workspace_bytes, void *stream);

/* ---- Multi-histogram reweighting (critmc/reweight.py:159-345) ------------ */

/* Bytes of device scratch for K runs and n_targets target betas. */
int64_t cm_mbar_workspace_bytes(int32_t K, int32_t n_targets);

/* MBAR objective of MultiHistogram._objective (critmc/reweight.py:231-247) over
 * U distinct shifted energies u_energy_dev [U] with log-multiplicities
 * u_logmult_dev [U]: out_dev [1 + K + K*K] = (F, gradient [K], Hessian [K][K])
 * at free energies f_dev [K]; betas_dev, log_n_dev (= ln N_k / g_k) [K];
 * 1 <= K <= 64.  ld_out_dev: null, or [U] receives ln D(E_u).  Partial sums
 * are combined in a fixed order, so results are reproducible. */
int cm_mbar_objective(const double *u_energy_dev, const double *u_logmult_dev, int64_t U,
                      const double *betas_dev, const double *log_n_dev, const double *f_dev,
                      int32_t K, double *out_dev, double *ld_out_dev, void *workspace_dev,
                      int64_t workspace_bytes, void *stream);

/* cm_mbar_objective without a workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight Newton loops in a serving process).  Same
 * kernels, launch shape and results as cm_mbar_objective; the per-workgroup
 * partials live in a static device array of 17,043,456 bytes per GPU (enough
 * for the largest objective, K = 64), held until the process ends.  K above
 * 64 is rejected with CM_EINVAL.  The array's device address is looked up on
 * the first call on each device (make that call before any graph capture).
 * Nothing in it needs clearing: each call writes every partial it reads.
 * All calls on one GPU share the array, so calls that overlap on that GPU
 * (on different streams, or from different threads) race; serialise them. */
int cm_mbar_objective_noalloc(const double *u_energy_dev, const double *u_logmult_dev,
                              int64_t U, const double *betas_dev, const double *log_n_dev,
                              const double *f_dev, int32_t K, double *out_dev,
                              double *ld_out_dev, void *stream);

/* ln sum_u exp(m_u - beta E_u - ld_u) for each of n_targets betas (MultiHistogram.log_z,
 * critmc/reweight.py:313-320, before the energy-shift correction). */
int cm_mbar_log_z(const double *u_energy_dev, const double *u_logmult_dev, const double *ld_dev,
                  int64_t U, const double *targets_dev, int32_t n_targets, double *out_dev,
                  void *workspace_dev, int64_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CRITMC_HIP_H */
