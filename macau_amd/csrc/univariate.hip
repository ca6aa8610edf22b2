// univariate.hip -- prior update of the univariate prior MacauOnePrior
// (macauoneprior.cpp:189-289).
//
//   sample_mu_lambda: the Normal-Wishart draw of the rows' U - Uhat (the
//     shared hyper kernels), of which only the diagonal lambda is kept;
//   sample_beta: for every latent d an independent chain over the features
//     in order f = 0..nfeat-1:
//       z = U_d - mu_d - Uhat_d,
//       A = lambda_beta_d + lambda_d colsq_f,  B = lambda_d (z . F_f + beta_df colsq_f),
//       beta_df ~ N(B / A, 1 / A),  z += F_f (beta_df_old - beta_df);
//   compute_uhat (the shared SpMM);
//   sample_lambda_beta: lambda_beta_d ~ Gamma(0.1 + nfeat / 2, 1 / (0.1 + |beta_d|^2 / 2)).
//
// The feature chain is sequential by construction, so it runs one wave per
// latent with z in LDS (rows up to kUniLdsRows; a global buffer beyond).
// Everything a step needs that does not depend on the chain -- 1 / A, the
// scaled normal sqrt(1/A) n_df, colsq_f, the old beta_df and the first 64
// entries of the next column -- is computed by a parallel pre-pass or
// prefetched one step ahead, leaving the LDS gather, one wave sum, three
// FP64 operations and the LDS scatter on the chain's critical path.
#include <math.h>

#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace mgpu {

// Lambda <- diag(Lambda) (FP64, K x K), Lambda_f <- diag (FP32, KP x KP, pad
// diagonal kept), Lmu_f = lambda * mu
__global__ void __launch_bounds__(256) k_uni_diag(int K, int KP, double* Lambda, float* Lambda_f, float* Lmu_f,
                                                  const double* mu) {
  for (int q = threadIdx.x; q < KP * KP; q += 256) {
    const int i = q % KP, j = q / KP;
    if (i != j) {
      Lambda_f[q] = 0.f;
      if (i < K && j < K) Lambda[i + j * K] = 0.0;
    }
  }
  for (int i = threadIdx.x; i < KP; i += 256) Lmu_f[i] = i < K ? (float)(Lambda[i + i * K] * mu[i]) : 0.f;
}

// per (d, f): ainv = 1 / A_df, nz = sqrt(ainv) n_df with n_df the normal
// (feature f, latent d, it, tag(BETA_ONE, mode)); layout [d * nfeat + f]
__global__ void __launch_bounds__(256) k_beta_uni_pre(BetaUniArgs a) {
  const int it = *a.iter;
  const uint32_t tag = make_tag(TAG_BETA_ONE, (uint32_t)a.mode);
  const size_t tot = (size_t)a.K * a.nfeat;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < tot; q += (size_t)gridDim.x * 256) {
    const int d = (int)(q / a.nfeat), f = (int)(q % a.nfeat);
    const double A = a.lambda_beta[d] + a.Lambda[d + d * a.K] * a.colsq[f];
    const double ainv = 1.0 / A;
    a.ainv[q] = ainv;
    a.nz[q] = sqrt(ainv) * normal_f64(a.seed, (uint32_t)f, (uint32_t)d, (uint32_t)it, tag);
  }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
  return x;
}

constexpr int kUniLdsRows = 19456;  // 152 KB of FP64 residuals

// z in LDS, or (LDS = false) in global memory accessed as volatile so that
// every gather sees the previous step's scatter by other lanes
template <bool LDS>
__global__ void __launch_bounds__(64) k_beta_uni_sweep(BetaUniArgs a) {
  using ZPtr = typename std::conditional<LDS, double*, volatile double*>::type;
  __shared__ double zs[LDS ? kUniLdsRows : 1];
  const int d = blockIdx.x, lane = threadIdx.x;
  const int K = a.K, N = a.N, nfeat = a.nfeat;
  ZPtr Z = LDS ? zs : a.Zw + (size_t)d * N;
  const double mud = a.mu[d], lamd = a.Lambda[d + d * K];
  for (int i = lane; i < N; i += 64) Z[i] = (double)a.U[(size_t)i * a.KP + d] - mud - a.uhat[(size_t)i * K + d];
  __syncthreads();
  const int* __restrict__ rp = a.Ft_rowptr;
  const int* __restrict__ cols = a.Ft_cols;
  const double* __restrict__ vals = a.Ft_vals;
  const int nnzF = rp[nfeat];
  const double* ainv = a.ainv + (size_t)d * nfeat;
  const double* nz = a.nz + (size_t)d * nfeat;
  // state of feature f, loaded one step ahead
  int beg = rp[0], end = nfeat > 0 ? rp[1] : 0;
  int c0 = 0;
  double v0 = 0.0;
  if (beg + lane < nnzF) {
    c0 = cols[beg + lane];
    v0 = vals ? vals[beg + lane] : 1.0;
  }
  double cs = nfeat > 0 ? a.colsq[0] : 0.0, bold = nfeat > 0 ? a.beta[d] : 0.0;
  double ai = nfeat > 0 ? ainv[0] : 0.0, nzf = nfeat > 0 ? nz[0] : 0.0;
  for (int f = 0; f < nfeat; f++) {
    // prefetch feature f + 1 (its column starts where this one ends)
    const bool more = f + 1 < nfeat;
    const int nend = more ? rp[f + 2] : end;
    int nc0 = 0;
    double nv0 = 0.0;
    if (end + lane < nnzF) {
      nc0 = cols[end + lane];
      nv0 = vals ? vals[end + lane] : 1.0;
    }
    const double ncs = more ? a.colsq[f + 1] : 0.0;
    const double nbold = more ? a.beta[(size_t)(f + 1) * K + d] : 0.0;
    const double nai = more ? ainv[f + 1] : 0.0, nnzf = more ? nz[f + 1] : 0.0;
    // zx = z . F_f
    double zx = 0.0;
    if (beg + lane < end) zx = v0 * Z[c0];
    for (int e = beg + 64 + lane; e < end; e += 64) zx += (vals ? vals[e] : 1.0) * Z[cols[e]];
    zx = wave_sum(zx);
    const double B = lamd * (zx + bold * cs);
    const double bnew = B * ai + nzf;
    const double delta = bold - bnew;
    if (lane == 0) a.beta[(size_t)f * K + d] = bnew;
    // z += F_f delta (the column's rows are distinct: no two lanes collide)
    if (beg + lane < end) Z[c0] += v0 * delta;
    for (int e = beg + 64 + lane; e < end; e += 64) Z[cols[e]] += (vals ? vals[e] : 1.0) * delta;
    __syncthreads();
    beg = end;
    end = nend;
    c0 = nc0;
    v0 = nv0;
    cs = ncs;
    bold = nbold;
    ai = nai;
    nzf = nnzf;
  }
}

// lambda_beta_d ~ Gamma(0.1 + nfeat / 2, scale 1 / (0.1 + sum_f beta_df^2 / 2)),
// the gamma keyed (d, it, tag(LAMBDA_BETA, mode)); one work-group per latent
__global__ void __launch_bounds__(256) k_lambda_beta_uni(BetaUniArgs a) {
  __shared__ double red[256];
  const int d = blockIdx.x;
  double s = 0.0;
  for (int f = threadIdx.x; f < a.nfeat; f += 256) {
    const double b = a.beta[(size_t)f * a.K + d];
    s += b * b;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double shape = 0.1 + (double)a.nfeat / 2.0, rate = 0.1 + red[0] / 2.0;
    const uint32_t tag = make_tag(TAG_LAMBDA_BETA, (uint32_t)a.mode);
    a.lambda_beta[d] = gamma_f64(a.seed, shape, (uint32_t)d, (uint32_t)*a.iter, tag) / rate;
  }
}

void launch_uni_diag(int K, int KP, double* Lambda, float* Lambda_f, float* Lmu_f, const double* mu, hipStream_t s) {
  hipLaunchKernelGGL(k_uni_diag, dim3(1), dim3(256), 0, s, K, KP, Lambda, Lambda_f, Lmu_f, mu);
}

bool beta_uni_in_lds(int N) { return N <= kUniLdsRows; }

void launch_beta_uni(const BetaUniArgs& a, hipStream_t s) {
  if (a.K <= 0) return;
  if (a.nfeat > 0) {
    size_t tot = (size_t)a.K * a.nfeat;
    size_t g = (tot + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_beta_uni_pre, dim3((unsigned)g), dim3(256), 0, s, a);
    if (beta_uni_in_lds(a.N))
      hipLaunchKernelGGL(k_beta_uni_sweep<true>, dim3(a.K), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL(k_beta_uni_sweep<false>, dim3(a.K), dim3(64), 0, s, a);
  }
}

void launch_lambda_beta_uni(const BetaUniArgs& a, hipStream_t s) {
  if (a.K > 0) hipLaunchKernelGGL(k_lambda_beta_uni, dim3(a.K), dim3(256), 0, s, a);
}

}  // namespace mgpu
