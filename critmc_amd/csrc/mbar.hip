// Multi-histogram (MBAR) free-energy objective on the device
// (critmc/reweight.py:159-345: MultiHistogram._objective and log_z).
//
// Runs k = 0..K-1 at inverse temperatures beta_k contribute samples whose
// distinct energies E_u carry log-multiplicities m_u (count times the run
// weight 1/g).  With log_n_k = ln(N_k / g_k),
//   a_uk  = log_n_k + f_k - beta_k E_u,   ld_u = ln sum_k exp(a_uk),
//   F     = sum_u e^{m_u} ld_u - sum_k e^{log_n_k} f_k,
//   w_uk  = exp(a_uk - ld_u),
//   grad_k = sum_u e^{m_u} w_uk - e^{log_n_k},
//   H_kl  = delta_kl sum_u e^{m_u} w_uk - sum_u e^{m_u} w_uk w_ul.
// One pass over the samples computes F, grad and H: each workgroup stages a
// tile of w_uk (times e^{m_u/2}) in LDS and accumulates the K x K product
// W^T W in registers; per-workgroup partials are summed in workgroup order by
// a second launch, so every result is reproducible run to run.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.hpp"

namespace cm {

constexpr int MBAR_KMAX = 64;
constexpr int MBAR_TILE = 64;                                 // samples per LDS tile
constexpr int MBAR_PAIRS = MBAR_KMAX * MBAR_KMAX / 256;       // (k, l) pairs per lane
constexpr int MBAR_GRID = 512;

// ln sum_k exp(log_n_k + f_k - beta_k e), two passes over k.
__device__ __forceinline__ double mbar_ld(double e, const double* lnf, const double* beta, int K)
{
    double mx = -INFINITY;
    for (int k = 0; k < K; ++k) mx = fmax(mx, lnf[k] - beta[k] * e);
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += exp(lnf[k] - beta[k] * e - mx);
    return mx + log(s);
}

__global__ __launch_bounds__(256) void mbar_partial(const double* __restrict__ E,
                                                    const double* __restrict__ logmult, int64_t U,
                                                    const double* __restrict__ betas,
                                                    const double* __restrict__ log_n,
                                                    const double* __restrict__ f, int K,
                                                    double* __restrict__ ld_out,
                                                    double* __restrict__ part)
{
    __shared__ double lnf[MBAR_KMAX], beta[MBAR_KMAX];
    __shared__ double W[MBAR_TILE][MBAR_KMAX + 1];  // sqrt(mult) * w, padded row
    __shared__ double red[256];
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        lnf[k] = log_n[k] + f[k];
        beta[k] = betas[k];
    }
    __syncthreads();
    const int nk2 = K * K;
    double h[MBAR_PAIRS];
#pragma unroll
    for (int p = 0; p < MBAR_PAIRS; ++p) h[p] = 0.0;
    double value = 0.0;
    const int64_t ntile = (U + MBAR_TILE - 1) / MBAR_TILE;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t u0 = tile * MBAR_TILE;
        // lanes 0..63: one sample each -> row of W
        if (threadIdx.x < MBAR_TILE) {
            const int64_t u = u0 + threadIdx.x;
            if (u < U) {
                const double e = E[u];
                const double ld = mbar_ld(e, lnf, beta, K);
                if (ld_out) ld_out[u] = ld;
                const double m = exp(logmult[u]);
                value += m * ld;
                const double sm = exp(0.5 * logmult[u]);
                for (int k = 0; k < K; ++k) W[threadIdx.x][k] = sm * exp(lnf[k] - beta[k] * e - ld);
            } else {
                for (int k = 0; k < K; ++k) W[threadIdx.x][k] = 0.0;
            }
        }
        __syncthreads();
        const int rows = (int)min<int64_t>(MBAR_TILE, U - u0);
#pragma unroll
        for (int p = 0; p < MBAR_PAIRS; ++p) {
            const int idx = threadIdx.x + 256 * p;
            if (idx < nk2) {
                const int k = idx / K, l = idx - (idx / K) * K;
                double acc = h[p];
                for (int r = 0; r < rows; ++r) acc += W[r][k] * W[r][l];
                h[p] = acc;
            }
        }
        __syncthreads();
    }
    // partial layout per workgroup: [value, diag-sum_k (K), WtW (K*K)]
    double* out = part + (int64_t)blockIdx.x * (1 + K + nk2);
#pragma unroll
    for (int p = 0; p < MBAR_PAIRS; ++p) {
        const int idx = threadIdx.x + 256 * p;
        if (idx < nk2) out[1 + K + idx] = h[p];
    }
    red[threadIdx.x] = value;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) v += red[t];
        out[0] = v;
    }
}

// Sums the partials in workgroup order; col_k = sum_l (W^T W)_kl because the
// weights of a sample sum to one over k.
__global__ void mbar_final(const double* __restrict__ part, int nblocks, int K,
                           const double* __restrict__ log_n, const double* __restrict__ f,
                           double* __restrict__ out)
{
    const int nk2 = K * K, stride = 1 + K + nk2;
    for (int idx = threadIdx.x; idx < nk2 + 1; idx += blockDim.x) {
        double s = 0.0;
        const int off = idx == 0 ? 0 : 1 + K + (idx - 1);
        for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * stride + off];
        if (idx == 0) out[0] = s;
        else out[1 + K + (idx - 1)] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = out[0];
        for (int k = 0; k < K; ++k) v -= exp(log_n[k]) * f[k];
        out[0] = v;
    }
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        double col = 0.0;
        for (int l = 0; l < K; ++l) col += out[1 + K + k * K + l];
        out[1 + k] = col - exp(log_n[k]);  // gradient
    }
    __syncthreads();
    // Hessian in place: diag(col) - W^T W
    for (int idx = threadIdx.x; idx < nk2; idx += blockDim.x) {
        const int k = idx / K, l = idx - (idx / K) * K;
        const double wtw = out[1 + K + idx];
        out[1 + K + idx] = (k == l ? out[1 + k] + exp(log_n[k]) : 0.0) - wtw;
    }
}

// ln sum_u exp(m_u - beta E_u - ld_u) for each target beta: per-workgroup
// (max, scaled sum) pairs, combined in order.
__global__ __launch_bounds__(256) void mbar_logz_partial(const double* __restrict__ E,
                                                         const double* __restrict__ logmult,
                                                         const double* __restrict__ ld, int64_t U,
                                                         const double* __restrict__ targets,
                                                         double* __restrict__ part)
{
    __shared__ double smx[256], ssum[256];
    const double b = targets[blockIdx.y];
    double mx = -INFINITY, s = 0.0;
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < U;
         u += (int64_t)gridDim.x * blockDim.x) {
        const double x = logmult[u] - b * E[u] - ld[u];
        if (x > mx) {
            s = s * exp(mx - x) + 1.0;
            mx = x;
        } else {
            s += exp(x - mx);
        }
    }
    smx[threadIdx.x] = mx;
    ssum[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double M = -INFINITY, S = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            if (smx[t] == -INFINITY) continue;
            if (smx[t] > M) {
                S = S * exp(M - smx[t]) + ssum[t];
                M = smx[t];
            } else {
                S += ssum[t] * exp(smx[t] - M);
            }
        }
        part[2 * ((int64_t)blockIdx.y * gridDim.x + blockIdx.x)] = M;
        part[2 * ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) + 1] = S;
    }
}

__global__ void mbar_logz_final(const double* __restrict__ part, int nblocks, int T,
                                double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double M = -INFINITY, S = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        const double m = part[2 * ((int64_t)t * nblocks + b)], s = part[2 * ((int64_t)t * nblocks + b) + 1];
        if (m == -INFINITY) continue;
        if (m > M) {
            S = S * exp(M - m) + s;
            M = m;
        } else {
            S += s * exp(m - M);
        }
    }
    out[t] = M + log(S);
}

}  // namespace cm

// Scratch of cm_mbar_objective_noalloc: the per-workgroup partials of the
// largest objective (K = 64 runs over 512 workgroups, 1 + K + K^2 doubles
// each), 17,043,456 bytes on every GPU the library runs on, held until the
// process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[(int64_t)cm::MBAR_GRID *
                                                (1 + cm::MBAR_KMAX + cm::MBAR_KMAX * cm::MBAR_KMAX)];

namespace cm {
namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kScratchDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int scratch_on_device(void** out)
{
    int dev = 0;
    CM_HIP_CHECK(hipGetDevice(&dev));
    CM_REQUIRE(dev >= 0 && dev < kScratchDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        CM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CM_REQUIRE(p != nullptr, "no device address for the MBAR scratch");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return CM_OK;
}

// Both objective entry points: partials into `part`, then the ordered sums.
int mbar_objective_launch(const double* u_energy, const double* u_logmult, int64_t U,
                          const double* betas, const double* log_n, const double* f, int K,
                          double* out, double* ld_out, double* part, hipStream_t st)
{
    const int64_t ntile = (U + MBAR_TILE - 1) / MBAR_TILE;
    const int grid = (int)std::min<int64_t>(ntile, MBAR_GRID);
    hipLaunchKernelGGL(mbar_partial, dim3(grid), dim3(256), 0, st, u_energy, u_logmult, U, betas,
                       log_n, f, K, ld_out, part);
    hipLaunchKernelGGL(mbar_final, dim3(1), dim3(256), 0, st, part, grid, K, log_n, f, out);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // namespace
}  // namespace cm

using namespace cm;

extern "C" {

int64_t cm_mbar_workspace_bytes(int32_t K, int32_t n_targets)
{
    const int64_t a = (int64_t)MBAR_GRID * (1 + K + (int64_t)K * K) * 8;
    const int64_t b = (int64_t)MBAR_GRID * std::max(n_targets, 1) * 16;
    return std::max(a, b);
}

int cm_mbar_objective(const double* u_energy, const double* u_logmult, int64_t U,
                      const double* betas, const double* log_n, const double* f, int32_t K,
                      double* out, double* ld_out, void* workspace, int64_t workspace_bytes,
                      void* stream)
{
    CM_REQUIRE(K >= 1 && K <= MBAR_KMAX, "MBAR takes 1..64 runs, got %d", K);
    CM_REQUIRE(U >= 1, "need at least one sample");
    CM_REQUIRE(workspace_bytes >= cm_mbar_workspace_bytes(K, 1), "workspace too small");
    return mbar_objective_launch(u_energy, u_logmult, U, betas, log_n, f, K, out, ld_out,
                                 (double*)workspace, (hipStream_t)stream);
}

int cm_mbar_objective_noalloc(const double* u_energy, const double* u_logmult, int64_t U,
                              const double* betas, const double* log_n, const double* f,
                              int32_t K, double* out, double* ld_out, void* stream)
{
    CM_REQUIRE(K >= 1 && K <= MBAR_KMAX,
               "MBAR without allocation takes 1..%d runs (its scratch is sized for that), got %d",
               MBAR_KMAX, K);
    CM_REQUIRE(U >= 1, "need at least one sample");
    CM_REQUIRE(cm_mbar_workspace_bytes(K, 1) <= (int64_t)sizeof(scratch),
               "objective needs %lld bytes of scratch, the static array holds %lld",
               (long long)cm_mbar_workspace_bytes(K, 1), (long long)sizeof(scratch));
    void* part = nullptr;
    if (int rc = scratch_on_device(&part)) return rc;
    return mbar_objective_launch(u_energy, u_logmult, U, betas, log_n, f, K, out, ld_out,
                                 (double*)part, (hipStream_t)stream);
}

int cm_mbar_log_z(const double* u_energy, const double* u_logmult, const double* ld, int64_t U,
                  const double* targets, int32_t n_targets, double* out, void* workspace,
                  int64_t workspace_bytes, void* stream)
{
    CM_REQUIRE(U >= 1 && n_targets >= 1 && n_targets <= 65535, "bad sizes");
    CM_REQUIRE(workspace_bytes >= cm_mbar_workspace_bytes(1, n_targets), "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    const int grid = (int)std::min<int64_t>((U + 255) / 256, MBAR_GRID);
    double* part = (double*)workspace;
    hipLaunchKernelGGL(mbar_logz_partial, dim3(grid, n_targets), dim3(256), 0, st, u_energy,
                       u_logmult, ld, U, targets, part);
    hipLaunchKernelGGL(mbar_logz_final, dim3((n_targets + 255) / 256), dim3(256), 0, st, part,
                       grid, n_targets, out);
    CM_HIP_CHECK(hipGetLastError());
    return CM_OK;
}

}  // extern "C"
