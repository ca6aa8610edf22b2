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
// The feature chain of a latent is sequential, so it runs on one wave with
// z in LDS (rows up to kUniLdsRows; a global buffer beyond).  Features whose
// columns share no row commute (updating one leaves every z entry the other
// reads unchanged), which the sweep uses twice without changing its result:
// features with an empty column are updated by the parallel pre-pass, and
// the host groups the others, in order, into batches: up to four
// consecutive short columns (at most 16 entries) with pairwise disjoint
// rows, 16 lanes each, or one long column on all 64 lanes.  Everything a
// step needs that does not depend on the chain is off its critical path:
// 1 / A and the scaled normal sqrt(1/A) n_df come from the pre-pass, and
// each lane loads its batch slot's values and first column entry kUniDepth
// batches ahead (the batch records 2 kUniDepth ahead).  A step is the LDS
// gather, a 16- or 64-lane DPP sum, three FP64 operations and the scatter.
#include <math.h>

#include <stdexcept>
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
    const double nzv = sqrt(ainv) * normal_f64(a.seed, (uint32_t)f, (uint32_t)d, (uint32_t)it, tag);
    a.ainv[q] = ainv;
    a.nz[q] = nzv;
    // an empty column couples to nothing: its update is B = 0, beta = nz,
    // done here; the sequential sweep walks the other features only
    if (a.Ft_rowptr[f + 1] == a.Ft_rowptr[f]) a.beta[(size_t)f * a.K + d] = nzv;
  }
}


constexpr int kUniLdsRows = 19456;  // 152 KB of FP64 residuals

// x moved across lanes by one DPP pattern (both 32-bit halves)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const int2 v = *reinterpret_cast<int2*>(&x);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, v.x, CTRL, 0xF, 0xF, false);
  r.y = __builtin_amdgcn_update_dpp(0, v.y, CTRL, 0xF, 0xF, false);
  return *reinterpret_cast<double*>(&r);
}

// sum over each 16-lane row in a fixed order (pairs, quads, half-row and
// row mirrors), returned in every lane of the row
__device__ __forceinline__ double row16_sum(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return x;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64_rows(double x) {
  const int2 v = *reinterpret_cast<int2*>(&x);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, v.x, CTRL, ROWMASK, 0xF, false);
  r.y = __builtin_amdgcn_update_dpp(0, v.y, CTRL, ROWMASK, 0xF, false);
  return *reinterpret_cast<double*>(&r);
}

// sum over the wave: the row sums, then rows combined by the row broadcasts
// into lane 63, which readlane distributes
__device__ __forceinline__ double wave_sum64(double x) {
  x = row16_sum(x);
  x += dpp_f64_rows<0x142, 0xA>(x);  // row_bcast15 into rows 1 and 3
  x += dpp_f64_rows<0x143, 0xC>(x);  // row_bcast31 into rows 2 and 3
  const int2 v = *reinterpret_cast<int2*>(&x);
  int2 r;
  r.x = __builtin_amdgcn_readlane(v.x, 63);
  r.y = __builtin_amdgcn_readlane(v.y, 63);
  return *reinterpret_cast<double*>(&r);
}

constexpr int kUniDepth = 4;     // batches whose loads are in flight
constexpr int kWideChunks = 16;  // 64-entry chunks of a long column loaded together

// Loads of one batch slot for this lane: one of its feature's first 16 column
// entries (64 for a long column; index clamped, masked at use), the
// feature's colsq, old beta, 1 / A and scaled normal.
struct UniSlot {
  int f, beg, end, lng;  // lng: the batch is one long column on all lanes
  int c;
  double v, cs, bo, ai, nz;
};

__device__ __forceinline__ UniSlot load_slot(const BetaUniArgs& a, int4 m, int d, int sub, int lane, int last) {
  UniSlot s;
  s.f = m.x;
  s.beg = m.y;
  s.end = m.z;
  s.lng = m.w;
  const int fc = max(m.x, 0);
  const int ix = min(m.y + (m.w ? lane : sub), last);
  s.c = a.Ft_cols[ix];
  s.v = a.Ft_vals[ix];
  s.cs = a.colsq[fc];
  s.bo = a.beta[(size_t)fc * a.K + d];
  s.ai = a.ainv[(size_t)d * a.nfeat + fc];
  s.nz = a.nz[(size_t)d * a.nfeat + fc];
  return s;
}

// z in LDS (one wave: its LDS operations complete in issue order, so a step
// sees the previous step's scatter without a barrier), or (LDS = false) in
// global memory accessed as volatile with a barrier after each scatter.
// Batches hold up to four consecutive features whose columns share no row
// (built on the host), so updating them in one step gives the sequential
// sweep's result: lanes 16 s .. 16 s + 15 handle slot s.
template <bool LDS>
__global__ void __launch_bounds__(64) k_beta_uni_sweep(BetaUniArgs a) {
  using ZPtr = typename std::conditional<LDS, double*, volatile double*>::type;
  __shared__ double zs[LDS ? kUniLdsRows : 1];
  const int d = blockIdx.x, lane = threadIdx.x, sub = lane & 15, slot = lane >> 4;
  const int K = a.K, N = a.N, nb = a.nbatch;
  ZPtr Z = LDS ? zs : a.Zw + (size_t)d * N;
  const double mud = a.mu[d], lamd = a.Lambda[d + d * K];
  for (int i = lane; i < N; i += 64) Z[i] = (double)a.U[(size_t)i * a.KP + d] - mud - a.uhat[(size_t)i * K + d];
  __syncthreads();
  if (nb <= 0) return;
  const int last = max(a.Ft_rowptr[a.nfeat] - 1, 0);  // the entry arrays hold at least one element
  const int4* __restrict__ meta = a.batch_meta;
  auto meta_of = [&](int b) { return meta[(size_t)min(b, nb - 1) * 4 + slot]; };
  // data ring: batch b in slot b % D; meta ring: the batch D ahead of it
  UniSlot dr[kUniDepth];
  int4 mr[kUniDepth];
#pragma unroll
  for (int q = 0; q < kUniDepth; q++) {
    dr[q] = load_slot(a, meta_of(q), d, sub, lane, last);
    mr[q] = meta_of(q + kUniDepth);
  }
  for (int b0 = 0; b0 < nb; b0 += kUniDepth) {
#pragma unroll
    for (int q = 0; q < kUniDepth; q++) {
      const int b = b0 + q;
      const UniSlot& s = dr[q];
      if (b < nb) {
        const bool wide = __builtin_amdgcn_readfirstlane(s.lng) != 0;  // uniform: one long column
        const int off = wide ? lane : sub;
        const bool live = s.beg + off < s.end;
        const double v0 = live ? s.v : 0.0;
        const double zc = live ? (double)Z[s.c] : 0.0;
        double zx = v0 * zc;
        if (wide) {
          // the rest of the column, kWideChunks 64-entry chunks per trip
          // from clamped indices so their loads overlap
          for (int e0 = s.beg + 64; e0 < s.end; e0 += 64 * kWideChunks) {
            int c4[kWideChunks];
            double v4[kWideChunks];
#pragma unroll
            for (int t = 0; t < kWideChunks; t++) {
              const int ix = min(e0 + 64 * t + lane, last);
              c4[t] = a.Ft_cols[ix];
              v4[t] = e0 + 64 * t + lane < s.end ? a.Ft_vals[ix] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < kWideChunks; t++) zx += v4[t] * Z[c4[t]];
          }
          zx = wave_sum64(zx);
        } else {
          zx = row16_sum(zx);
        }
        const double B = lamd * (zx + s.bo * s.cs);
        const double bnew = B * s.ai + s.nz;
        const double delta = s.bo - bnew;
        if (off == 0 && s.f >= 0) a.beta[(size_t)s.f * K + d] = bnew;
        // z += F_f delta (distinct rows within and across the batch's columns)
        if (live) Z[s.c] = zc + v0 * delta;
        if (wide) {
          for (int e0 = s.beg + 64; e0 < s.end; e0 += 64 * kWideChunks) {
            int c4[kWideChunks];
            double v4[kWideChunks];
#pragma unroll
            for (int t = 0; t < kWideChunks; t++) {
              const int ix = min(e0 + 64 * t + lane, last);
              c4[t] = a.Ft_cols[ix];
              v4[t] = a.Ft_vals[ix];
            }
#pragma unroll
            for (int t = 0; t < kWideChunks; t++)
              if (e0 + 64 * t + lane < s.end) Z[c4[t]] += v4[t] * delta;
          }
        }
        if (LDS) {
          __builtin_amdgcn_wave_barrier();
        } else {
          __syncthreads();
        }
      }
      // refills once the slot's values are dead, unconditionally so the
      // loads write the ring registers in place
      dr[q] = load_slot(a, mr[q], d, sub, lane, last);
      mr[q] = meta_of(b + 2 * kUniDepth);
    }
  }
}

// sum_f beta_df^2 per latent d in one coalesced pass: work-group b takes a
// contiguous range of features; thread t < K (256 / K) holds latent
// t % K and walks the features t / K, t / K + 256 / K, ... of the range
// (consecutive threads read consecutive entries of beta); the threads of a
// latent are combined in thread order into partial[b][d], and the blocks in
// block order by launch_partial_reduce
constexpr int kLbBlocks = 64;
__global__ void __launch_bounds__(256) k_beta_sumsq(BetaUniArgs a, double* partial) {
  __shared__ double red[256];
  const int K = a.K, per = 256 / K, t = threadIdx.x;
  const int chunk = (a.nfeat + gridDim.x - 1) / gridDim.x;
  const int f0 = blockIdx.x * chunk, f1 = min(a.nfeat, f0 + chunk);
  double s = 0.0;
  if (t < K * per) {
    const int d = t % K;
    for (int f = f0 + t / K; f < f1; f += per) {
      const double b = a.beta[(size_t)f * K + d];
      s += b * b;
    }
  }
  red[t] = s;
  __syncthreads();
  if (t < K) {
    double v = 0.0;
    for (int m = 0; m < per; m++) v += red[t + m * K];
    partial[(size_t)blockIdx.x * K + t] = v;
  }
}

// lambda_beta_d ~ Gamma(0.1 + nfeat / 2, scale 1 / (0.1 + sum_f beta_df^2 / 2)),
// the gamma keyed (d, it, tag(LAMBDA_BETA, mode))
__global__ void __launch_bounds__(256) k_lambda_beta_uni(BetaUniArgs a, const double* sumsq) {
  for (int d = threadIdx.x; d < a.K; d += 256) {
    const double shape = 0.1 + (double)a.nfeat / 2.0, rate = 0.1 + sumsq[d] / 2.0;
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

int lambda_beta_uni_partials(int K) { return kLbBlocks * K; }

void launch_lambda_beta_uni(const BetaUniArgs& a, double* partial, hipStream_t s) {
  if (a.K <= 0) return;
  if (a.K > 256) throw std::runtime_error("lambda_beta: more than 256 latent dimensions");
  hipLaunchKernelGGL(k_beta_sumsq, dim3(kLbBlocks), dim3(256), 0, s, a, partial);
  launch_partial_reduce(partial, kLbBlocks, a.K, s);
  hipLaunchKernelGGL(k_lambda_beta_uni, dim3(1), dim3(256), 0, s, a, (const double*)partial);
}

}  // namespace mgpu
