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
// Everything a step needs that does not depend on the chain is off its
// critical path: 1 / A and the scaled normal sqrt(1/A) n_df come from a
// parallel pre-pass; the per-feature values of 64 features are loaded one
// per lane, a chunk ahead, and read with readlane; the first 64 entries of
// each column are loaded kUniRing features ahead.  A step is then the LDS
// gather, one wave sum, three FP64 operations and the LDS scatter.
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


constexpr int kUniLdsRows = 19456;  // 152 KB of FP64 residuals

// Per-feature values of 64 consecutive features, lane j holding feature
// f0 + j (zero past the end).
struct UniChunk {
  int rb, re;                  // column extent in F^T
  double cs, bo, ai, nz;       // colsq, old beta, 1 / A, sqrt(1 / A) n
};

// unconditional loads from clamped indices (features past the end repeat
// the last one and are never used), so no load sits under a lane mask
__device__ __forceinline__ UniChunk load_chunk(const BetaUniArgs& a, int d, int f0, int lane) {
  UniChunk c;
  const int f = min(f0 + lane, a.nfeat - 1);
  c.rb = a.Ft_rowptr[f];
  c.re = a.Ft_rowptr[f + 1];
  c.cs = a.colsq[f];
  c.bo = a.beta[(size_t)f * a.K + d];
  c.ai = a.ainv[(size_t)d * a.nfeat + f];
  c.nz = a.nz[(size_t)d * a.nfeat + f];
  return c;
}

__device__ __forceinline__ int rdl(int x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ __forceinline__ double rdl(double x, int j) {
  const int2 v = *reinterpret_cast<int2*>(&x);
  int2 r;
  r.x = __builtin_amdgcn_readlane(v.x, j);
  r.y = __builtin_amdgcn_readlane(v.y, j);
  return *reinterpret_cast<double*>(&r);
}

// x moved across lanes by one DPP pattern (both 32-bit halves); lanes
// outside ROWMASK receive 0
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double x) {
  const int2 v = *reinterpret_cast<int2*>(&x);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, v.x, CTRL, ROWMASK, 0xF, false);
  r.y = __builtin_amdgcn_update_dpp(0, v.y, CTRL, ROWMASK, 0xF, false);
  return *reinterpret_cast<double*>(&r);
}

// sum over the wave in a fixed order, returned in every lane: pairs and
// quads (quad_perm), 8 and 16 lanes (half-row and row mirror), then rows
// combined by the row broadcasts into lane 63, which readlane distributes
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141, 0xF>(x);  // row_half_mirror
  x += dpp_f64<0x140, 0xF>(x);  // row_mirror
  x += dpp_f64<0x142, 0xA>(x);  // row_bcast15 into rows 1 and 3
  x += dpp_f64<0x143, 0xC>(x);  // row_bcast31 into rows 2 and 3
  return rdl(x, 63);
}

constexpr int kUniRing = 8;  // columns whose first 64 entries are in flight

// z in LDS (one wave: its LDS operations complete in issue order, so a step
// sees the previous step's scatter without a barrier), or (LDS = false) in
// global memory accessed as volatile with a barrier after each scatter
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
  const int* __restrict__ cols = a.Ft_cols;
  const double* __restrict__ vals = a.Ft_vals;

  if (nfeat <= 0) return;
  const int last = max(a.Ft_rowptr[nfeat] - 1, 0);  // the entry arrays hold at least one element
  UniChunk cur = load_chunk(a, d, 0, lane);
  UniChunk nxt = load_chunk(a, d, 64, lane);
  // ring slot q holds the raw loads at min(column start + lane, last) of the
  // feature f with f % kUniRing == q; lanes past the column's end are
  // masked when the slot is used
  int rc[kUniRing];
  double rv[kUniRing];
#pragma unroll
  for (int q = 0; q < kUniRing; q++) {
    const int ix = min(rdl(cur.rb, q) + lane, last);
    rc[q] = cols[ix];
    rv[q] = vals[ix];
  }
  for (int f0 = 0; f0 < nfeat; f0 += 64) {
    double bnew_mine = cur.bo;
    const int nloc = nfeat - f0 < 64 ? nfeat - f0 : 64;
    for (int j0 = 0; j0 < nloc; j0 += kUniRing) {
#pragma unroll
      for (int q = 0; q < kUniRing; q++) {
        const int j = j0 + q;
        const int c0 = rc[q];
        const double v0r = rv[q];
        // refill the slot with feature f0 + j + kUniRing (this chunk or the
        // next), unconditionally so the load writes the slot's register
        {
          const int jn = j + kUniRing;
          const int bc = rdl(cur.rb, jn & 63), bn = rdl(nxt.rb, jn & 63);
          const int ix = min((jn < 64 ? bc : bn) + lane, last);
          rc[q] = cols[ix];
          rv[q] = vals[ix];
        }
        if (j < nloc) {
          const int beg = rdl(cur.rb, j), end = rdl(cur.re, j);
          const bool live = beg + lane < end;
          const double v0 = live ? v0r : 0.0;
          const double zc = live ? (double)Z[c0] : 0.0;
          double zx = v0 * zc;
          const bool longcol = end - beg > 64;  // uniform
          if (longcol)
            for (int e = beg + 64 + lane; e < end; e += 64) zx += vals[e] * Z[cols[e]];
          zx = wave_sum(zx);
          const double bold = rdl(cur.bo, j), cs = rdl(cur.cs, j);
          const double B = lamd * (zx + bold * cs);
          const double bnew = B * rdl(cur.ai, j) + rdl(cur.nz, j);
          const double delta = bold - bnew;
          if (lane == j) bnew_mine = bnew;
          // z += F_f delta (the column's rows are distinct: no two lanes collide)
          if (live) Z[c0] = zc + v0 * delta;
          if (longcol)
            for (int e = beg + 64 + lane; e < end; e += 64) Z[cols[e]] += vals[e] * delta;
          if (LDS) {
            __builtin_amdgcn_wave_barrier();
          } else {
            __syncthreads();
          }
        }
      }
    }
    if (lane < nloc) a.beta[(size_t)(f0 + lane) * K + d] = bnew_mine;
    cur = nxt;
    nxt = load_chunk(a, d, f0 + 128, lane);
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
