// densechol.hip -- the F^T F branch of MacauPrior::sample_beta
// (latentprior.cpp:280-289): beta = Ft_y (F^T F + lambda_beta I)^{-1}.
//
// F^T F (n x n, n = number of features, lower triangle, FP64, col-major
// ld = n) is formed once at init (At_mul_A, linop.cpp:93-118).  Each Gibbs
// iteration copies it with lambda_beta on the diagonal into a work matrix W,
// factors W = L L^T in place (LAPACK dpotrf semantics, chol.cpp:13-31) and
// solves L L^T X = B for the K right-hand sides (chol_solve_t,
// chol.cpp:47-53).  Everything is blocked by 32 columns:
//   factor, block k:  k_potrf_diag (one wave, the 32 x 32 diagonal block),
//                     k_trsm_panel (one thread per row below it),
//                     k_syrk_tile  (one work-group per 32 x 32 lower tile of
//                                   the trailing matrix, panel tiles in LDS);
//   solves, block k:  k_trsv_block (one thread per right-hand side),
//                     k_fwd_update / k_bwd_update (one thread per output).
// B and X are n x nrhs row-major (feature f's nrhs values contiguous), i.e.
// the K x nfeat column-major layout of Ft_y and beta.
#include <stdexcept>

#include "common.h"
#include "kernels.h"

namespace mgpu {

constexpr int kCB = 32;            // block size
constexpr int kCLd = kCB + 1;      // LDS row stride of a 32 x 32 tile

// thread f1 owns column f1 of the lower triangle and accumulates, for every
// row of F holding f1 (in F^T order), every f2 >= f1 of that row -- the
// reference's loop order, so the sums are bit-identical to it
__global__ void k_ftf(CsrDev F, CsrDev Ft, double* out) {
  const int f1 = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = F.ncol;
  if (f1 >= n) return;
  double* col = out + (size_t)f1 * n;
  for (int i = Ft.row_ptr[f1]; i < Ft.row_ptr[f1 + 1]; i++) {
    const int r = Ft.cols[i];
    const double v1 = Ft.vals ? Ft.vals[i] : 1.0;
    for (int j = F.row_ptr[r]; j < F.row_ptr[r + 1]; j++) {
      const int f2 = F.cols[j];
      if (f1 <= f2) col[f2] += (F.vals ? F.vals[j] : 1.0) * v1;
    }
  }
}

// W = lower(FtF) + reg I (the upper triangle of W is never read)
__global__ void k_ftf_shift(const double* __restrict__ FtF, int n, const double* reg, double* __restrict__ W) {
  const double r = *reg;
  const size_t tot = (size_t)n * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % n), j = (int)(q / n);
    if (i >= j) W[q] = FtF[q] + (i == j ? r : 0.0);
  }
}

// the nb x nb diagonal block at k0 (nb <= 32) factored in place by one wave,
// lane i = row i, left-looking: s_i = A[i][k] - sum_{j<k} L[i][j] L[k][j],
// L[k][k] = sqrt(s_k), L[i][k] = s_i / L[k][k].  A non-positive pivot sets
// *fail (dpotrf info > 0) and is replaced by 1 so the rest stays finite.
__global__ void __launch_bounds__(64) k_potrf_diag(double* W, int n, int k0, int* fail) {
  __shared__ double A[kCB][kCLd];
  const int nb = min(kCB, n - k0);
  const int lane = threadIdx.x;
  const int i = lane < nb ? lane : nb - 1;  // lanes >= nb shadow the last row
  {
    // all 32 loads issued before the first LDS store
    double v[kCB];
#pragma unroll
    for (int j = 0; j < kCB; j++) v[j] = (j < nb && j <= i) ? W[(size_t)(k0 + i) + (size_t)(k0 + j) * n] : 0.0;
#pragma unroll
    for (int j = 0; j < kCB; j++) A[i][j] = v[j];
  }
  lds_wait();
  bool bad = false;
  for (int k = 0; k < nb; k++) {
    double s = A[i][k];
#pragma unroll 8
    for (int j = 0; j < k; j++) s -= A[i][j] * A[k][j];
    const double d = __shfl(s, k, 64);
    bad = bad || !(d > 0.0);
    const double lkk = sqrt(d > 0.0 ? d : 1.0);
    if (i == k) A[i][k] = lkk;
    else if (i > k) A[i][k] = s / lkk;
    lds_wait();
  }
  if (lane < nb)
    for (int j = 0; j <= i; j++) W[(size_t)(k0 + i) + (size_t)(k0 + j) * n] = A[i][j];
  if (lane == 0 && bad) atomicMax(fail, kFailDenseChol);
}

// loads the 32 x 32 lower block at (k0, k0) into L with all of a thread's
// loads issued before its LDS stores (T threads, kCB * kCB / T loads each)
template <int T>
__device__ __forceinline__ void load_diag_block(const double* W, int n, int k0, int nb, double (*L)[kCLd]) {
  constexpr int PER = kCB * kCB / T;
  double v[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int q = threadIdx.x + T * u;
    const int r = q / kCB, c = q % kCB;
    v[u] = (r < nb && c <= r) ? W[(size_t)(k0 + r) + (size_t)(k0 + c) * n] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int q = threadIdx.x + T * u;
    L[q / kCB][q % kCB] = v[u];
  }
}

// rows i >= k0 + 32 of the panel: w_i <- w_i L_kk^{-T} (forward substitution
// over the 32 columns, one thread per row)
__global__ void __launch_bounds__(256) k_trsm_panel(double* W, int n, int k0) {
  __shared__ double L[kCB][kCLd];
  load_diag_block<256>(W, n, k0, kCB, L);
  __syncthreads();
  const int i = k0 + kCB + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[kCB];
#pragma unroll
  for (int t = 0; t < kCB; t++) x[t] = W[(size_t)i + (size_t)(k0 + t) * n];
#pragma unroll
  for (int t = 0; t < kCB; t++) {
    double s = x[t];
#pragma unroll
    for (int u = 0; u < t; u++) s -= x[u] * L[t][u];
    x[t] = s / L[t][t];
  }
#pragma unroll
  for (int t = 0; t < kCB; t++) W[(size_t)i + (size_t)(k0 + t) * n] = x[t];
}

// trailing update of the lower tiles: W[i][j] -= sum_t W[i][k0+t] W[j][k0+t]
// for k0 + 32 <= j <= i < n; one work-group per 32 x 32 tile (tile row ti >=
// tile column tj of the trailing matrix), 4 outputs per thread
__global__ void __launch_bounds__(256) k_syrk_tile(double* W, int n, int k0) {
  __shared__ double Pi[kCB][kCLd], Pj[kCB][kCLd];
  const int b = blockIdx.x;
  int ti = (int)((sqrtf(8.f * (float)b + 1.f) - 1.f) * 0.5f);
  while ((ti + 1) * (ti + 2) / 2 <= b) ti++;
  while (ti * (ti + 1) / 2 > b) ti--;
  const int tj = b - ti * (ti + 1) / 2;
  const int k1 = k0 + kCB;
  const int i0 = k1 + ti * kCB, j0 = k1 + tj * kCB;
  {
    // all of a thread's loads issued before its LDS stores
    constexpr int PER = kCB * kCB / 256;
    double vi[PER], vj[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int q = threadIdx.x + 256 * u;
      const int r = q % kCB, t = q / kCB;  // consecutive threads: consecutive rows
      vi[u] = i0 + r < n ? W[(size_t)(i0 + r) + (size_t)(k0 + t) * n] : 0.0;
      vj[u] = j0 + r < n ? W[(size_t)(j0 + r) + (size_t)(k0 + t) * n] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int q = threadIdx.x + 256 * u;
      Pi[q % kCB][q / kCB] = vi[u];
      Pj[q % kCB][q / kCB] = vj[u];
    }
  }
  __syncthreads();
  const int r = threadIdx.x % kCB, c0 = (threadIdx.x / kCB) * 4;  // row r, columns c0..c0+3
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int t = 0; t < kCB; t++) {
    const double a = Pi[r][t];
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u] += a * Pj[c0 + u][t];
  }
  const int i = i0 + r;
  if (i >= n) return;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int j = j0 + c0 + u;
    if (j < n && j <= i) W[(size_t)i + (size_t)j * n] -= acc[u];
  }
}

// the nb rows of block k0 for every right-hand side (thread = rhs column):
// forward y_t = (x_t - sum_{s<t} L[t][s] y_s) / L[t][t], or backward
// x_t = (y_t - sum_{s>t} L[s][t] x_s) / L[t][t]
template <bool FWD>
__global__ void __launch_bounds__(128) k_trsv_block(const double* W, int n, int k0, double* X, int nrhs) {
  __shared__ double L[kCB][kCLd];
  const int nb = min(kCB, n - k0);
  load_diag_block<128>(W, n, k0, nb, L);
  __syncthreads();
  const int k = threadIdx.x;
  if (k >= nrhs) return;
  double x[kCB];
#pragma unroll
  for (int t = 0; t < kCB; t++) x[t] = t < nb ? X[(size_t)(k0 + t) * nrhs + k] : 0.0;
  if (FWD) {
#pragma unroll
    for (int t = 0; t < kCB; t++) {
      if (t < nb) {
        double s = x[t];
#pragma unroll
        for (int u = 0; u < t; u++) s -= L[t][u] * x[u];
        x[t] = s / L[t][t];
      }
    }
  } else {
#pragma unroll
    for (int t = kCB - 1; t >= 0; t--) {
      if (t < nb) {
        double s = x[t];
#pragma unroll
        for (int u = t + 1; u < kCB; u++)
          if (u < nb) s -= L[u][t] * x[u];
        x[t] = s / L[t][t];
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kCB; t++)
    if (t < nb) X[(size_t)(k0 + t) * nrhs + k] = x[t];
}

// rows i >= k0 + 32: X[i][k] -= sum_t L[i][k0+t] Y[k0+t][k]
__global__ void __launch_bounds__(256) k_fwd_update(const double* W, int n, int k0, double* X, int nrhs) {
  const size_t tot = (size_t)(n - k0 - kCB) * nrhs;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = k0 + kCB + (int)(q / nrhs), k = (int)(q % nrhs);
    double s = 0.0;
#pragma unroll 8
    for (int t = 0; t < kCB; t++) s += W[(size_t)i + (size_t)(k0 + t) * n] * X[(size_t)(k0 + t) * nrhs + k];
    X[(size_t)i * nrhs + k] -= s;
  }
}

// rows i < k0: X[i][k] -= sum_{t<nb} L[k0+t][i] X[k0+t][k]
__global__ void __launch_bounds__(256) k_bwd_update(const double* W, int n, int k0, double* X, int nrhs) {
  const int nb = min(kCB, n - k0);
  const size_t tot = (size_t)k0 * nrhs;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q / nrhs), k = (int)(q % nrhs);
    const double* li = W + (size_t)k0 + (size_t)i * n;  // L[k0 + t][i], t = 0 .. nb-1, contiguous
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < kCB; t++)
      if (t < nb) s += li[t] * X[(size_t)(k0 + t) * nrhs + k];
    X[(size_t)i * nrhs + k] -= s;
  }
}

static int grid_of(size_t work, int per) {
  size_t g = (work + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (int)g;
}

void launch_ftf(const CsrDev& F, const CsrDev& Ft, double* FtF, hipStream_t s) {
  const int n = F.ncol;
  if (hipMemsetAsync(FtF, 0, (size_t)n * n * sizeof(double), s) != hipSuccess)
    throw std::runtime_error("F^T F: clearing the output failed");
  hipLaunchKernelGGL(k_ftf, dim3((n + 127) / 128), dim3(128), 0, s, F, Ft, FtF);
}

void launch_ftf_solve(const double* FtF, int n, const double* reg, double* W, double* X, int nrhs, int* fail,
                      hipStream_t s) {
  if (nrhs > 128) throw std::runtime_error("F^T F solve: at most 128 right-hand sides");
  hipLaunchKernelGGL(k_ftf_shift, dim3(grid_of((size_t)n * n, 256)), dim3(256), 0, s, FtF, n, reg, W);
  for (int k0 = 0; k0 < n; k0 += kCB) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, s, W, n, k0, fail);
    const int rest = n - k0 - kCB;
    if (rest > 0) {
      hipLaunchKernelGGL(k_trsm_panel, dim3((rest + 255) / 256), dim3(256), 0, s, W, n, k0);
      const int T = (rest + kCB - 1) / kCB;
      hipLaunchKernelGGL(k_syrk_tile, dim3(T * (T + 1) / 2), dim3(256), 0, s, W, n, k0);
    }
  }
  for (int k0 = 0; k0 < n; k0 += kCB) {
    hipLaunchKernelGGL((k_trsv_block<true>), dim3(1), dim3(128), 0, s, (const double*)W, n, k0, X, nrhs);
    const int rest = n - k0 - kCB;
    if (rest > 0)
      hipLaunchKernelGGL(k_fwd_update, dim3(grid_of((size_t)rest * nrhs, 256)), dim3(256), 0, s, (const double*)W, n,
                         k0, X, nrhs);
  }
  for (int k0 = ((n - 1) / kCB) * kCB; k0 >= 0; k0 -= kCB) {
    hipLaunchKernelGGL((k_trsv_block<false>), dim3(1), dim3(128), 0, s, (const double*)W, n, k0, X, nrhs);
    if (k0 > 0)
      hipLaunchKernelGGL(k_bwd_update, dim3(grid_of((size_t)k0 * nrhs, 256)), dim3(256), 0, s, (const double*)W, n,
                         k0, X, nrhs);
  }
}

}  // namespace mgpu
