// sampler_plain.hip -- plain FP32 per-row latent sampler.
//
// Computes, for every row n of one mode (sample_latent_blas,
// latentprior.cpp:361-385; sample_latent_tensor, latentprior.cpp:118-161):
//   P = Lambda + alpha * sum_e c_e c_e^T,   b = alpha * sum_e c_e r_e + Lambda mu_n
//   P = L L^T,   u_n = L^{-T} (L^{-1} b + z_n)
// where c_e is the (Hadamard product of the) other modes' latent columns at
// nonzero e and r_e = y_e - mean.
//
// Layout: one work-group handles G = 64/KP rows (KP <= 64) or one row
// (KP = 128); a group of KP lanes owns one row and lane i owns row i of P
// in registers.  The Gram is accumulated in FP32 on the vector ALUs from
// nonzeros staged through LDS, in segments of 2048 nonzeros whose sums are
// added in order.  The Cholesky is right-looking with the
// current column broadcast through LDS; the forward solve L y = b rides
// along as an extra column; the back solve reads L row-major from LDS.
// Rows are processed in natural order with no load balancing: this is the
// plain baseline the batched sampler is measured against.
//
// Univariate variant (MacauOnePrior::sample_latents, macauoneprior.cpp:42-95
// and 108-187): the reference updates u_d for d = 0..K-1 in turn,
//   u_d ~ N(L_d / Q_d, 1 / Q_d),  Q_d = lambda_d + alpha sum_e c_ed^2,
//   L_d = lambda_d m_d + alpha sum_e (y_e - mean - sum_{j != d} u_j c_ej) c_ed,
// with u_j already updated for j < d.  With P and b assembled as above
// (Lambda = diag(lambda), b = h) this is one Gauss-Seidel sweep from the
// row's current value: u_d = (h_d - sum_{j != d} P_dj u_j) / P_dd +
// z_d / sqrt(P_dd).  Lane j holds row j of P and its own u_j, so step d is
// one lane-group sum of P_jd u_j and an update on lane d.
#include "common.h"
#include "kernels.h"

namespace mgpu {

// NO other modes (NO = kMaxOther: no = the run-time count, at most NO)
template <int NO>
__device__ __forceinline__ float gather_comp(const FactorSet& F, const int* __restrict__ idx, int e, int i,
                                             int KP, int no) {
  float x = F.f[0][(size_t)idx[(size_t)e * no] * KP + i];
#pragma unroll
  for (int t = 1; t < NO; t++)
    if (t < no) x *= F.f[t][(size_t)idx[(size_t)e * no + t] * KP + i];
  return x;
}

template <int KP, int NO, bool UNI>
__global__ void __launch_bounds__(KP < 64 ? 64 : KP) k_sample_plain(SampleArgs a) {
  constexpr int WG = KP < 64 ? 64 : KP;
  constexpr int G = WG / KP;
  constexpr int CH = 16;
  __shared__ float vbuf[G][CH][KP];
  __shared__ float rbuf[G][CH];
  __shared__ float colb[2][G][KP + 1];
  __shared__ float Lrow[G][KP][KP + 1];
  __shared__ float xs[2][G];

  const int tid = threadIdx.x, g = tid / KP, i = tid % KP;
  const int row = blockIdx.x * G + g;
  const bool active = row < a.rows.N;
  int beg = 0, end = 0;
  if (active) {
    beg = a.rows.row_ptr[row];
    end = a.rows.row_ptr[row + 1];
  }
  const int nnz = end - beg;
  const int it = *a.iter;

  // ---- Gram and right-hand side ----
  // acc sums at most kSeg nonzeros; each full segment is then added into this
  // lane's row of Lrow (a lane-private LDS row until the back solve), so a row
  // of ~10^6 nonzeros is summed as ~500 FP32 segment sums instead of one
  // running FP32 sum
  constexpr int kSegSteps = 2048 / CH;
  float acc[KP];
#pragma unroll
  for (int j = 0; j < KP; j++) {
    acc[j] = 0.f;
    Lrow[g][i][j] = 0.f;
  }
  float bi = 0.f, bseg = 0.f;
  int steps = 0;
  for (int base = 0; __syncthreads_or(base < nnz); base += CH) {
    if (steps == kSegSteps) {
#pragma unroll
      for (int j = 0; j < KP; j++) {
        Lrow[g][i][j] += acc[j];
        acc[j] = 0.f;
      }
      bi += bseg;
      bseg = 0.f;
      steps = 0;
    }
    steps++;
#pragma unroll
    for (int c = 0; c < CH; c++) {
      float x = 0.f, rr = 0.f;
      if (base + c < nnz) {
        const int e = beg + base + c;
        x = gather_comp<NO>(a.others, a.rows.idx, e, i, KP, NO < kMaxOther ? NO : a.rows.nother);
        rr = a.rows.r[e];
      }
      vbuf[g][c][i] = x;
      if (i == 0) rbuf[g][c] = rr;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CH; c++) {
      const float vi = vbuf[g][c][i];
      const float rc = rbuf[g][c];
#pragma unroll
      for (int j = 0; j < KP; j++) acc[j] += vi * vbuf[g][c][j];
      bseg += vi * rc;
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < KP; j++) acc[j] += Lrow[g][i][j];
  bi += bseg;

  // ---- P = Lambda + alpha G,  b = alpha c + Lambda mu ----
  const float alpha = (float)(*a.alpha);
#pragma unroll
  for (int j = 0; j < KP; j++) acc[j] = a.Lambda_f[i * KP + j] + alpha * acc[j];
  float b = alpha * bi;
  if (a.Lmu_rows) {
    if (active) b += a.Lmu_rows[(size_t)row * KP + i];
  } else {
    b += a.Lmu_f[i];
  }

  if constexpr (UNI) {
    // ---- one Gauss-Seidel sweep from the current value ----
    __shared__ float part[2][2];
    float u = active ? a.U[(size_t)row * KP + i] : 0.f;
    const float z =
        i < a.K ? normal_f32(a.seed, (uint32_t)row, (uint32_t)i, (uint32_t)it, make_tag(TAG_LATENT, a.mode)) : 0.f;
    if constexpr (KP <= 64) {
      // the KP-lane group is inside one wave: P's row in registers
#pragma unroll
      for (int d = 0; d < KP; d++) {
        if (d >= a.K) continue;  // uniform across the work-group
        float t = acc[d] * u;  // P[i][d] u_i = P[d][i] u_i
#pragma unroll
        for (int off = 1; off < KP; off <<= 1) t += __shfl_xor(t, off, KP);
        if (i == d) {
          float pdd = acc[d];
          if (!(pdd > 0.f)) {
            if (active) atomicMax(a.fail, row + 1);
            pdd = 1.f;
          }
          u = (b - t + pdd * u) / pdd + z / sqrtf(pdd);
        }
      }
    } else {
      // two waves per row: P's rows staged in LDS, wave partials combined
      // through LDS (double-buffered, one barrier per step)
#pragma unroll
      for (int j = 0; j < KP; j++) Lrow[0][i][j] = acc[j];
      __syncthreads();
      for (int d = 0; d < a.K; d++) {
        const float pid = Lrow[0][i][d];
        float t = pid * u;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) t += __shfl_xor(t, off, 64);
        if ((tid & 63) == 0) part[d & 1][tid >> 6] = t;
        __syncthreads();
        t = part[d & 1][0] + part[d & 1][1];
        if (i == d) {
          float pdd = pid;
          if (!(pdd > 0.f)) {
            if (active) atomicMax(a.fail, row + 1);
            pdd = 1.f;
          }
          u = (b - t + pdd * u) / pdd + z / sqrtf(pdd);
        }
      }
    }
    if (active) a.U[(size_t)row * KP + i] = u;
    return;
  }

  // ---- Cholesky with fused forward solve ----
#pragma unroll
  for (int k = 0; k < KP; k++) {
    float* cb = colb[k & 1][g];
    if (i >= k) cb[i] = acc[k];
    if (i == k) cb[KP] = b;
    __syncthreads();
    float dkk = cb[k];
    if (!(dkk > 0.f)) {
      if (i == k && active) atomicMax(a.fail, row + 1);
      dkk = 1.f;
    }
    const float sq = sqrtf(dkk);
    const float bk = cb[KP];
    if (i > k) {
      const float s = acc[k] / dkk;
#pragma unroll
      for (int j = k + 1; j < KP; j++) acc[j] -= s * cb[j];
      b -= s * bk;
      acc[k] = acc[k] / sq;
    } else if (i == k) {
      acc[k] = sq;
      b = bk / sq;
    }
  }

  // ---- add z, back solve L^T x = y ----
  if (i < a.K) b += normal_f32(a.seed, (uint32_t)row, (uint32_t)i, (uint32_t)it, make_tag(TAG_LATENT, a.mode));
#pragma unroll
  for (int j = 0; j < KP; j++) Lrow[g][i][j] = (j <= i) ? acc[j] : 0.f;
  __syncthreads();
#pragma unroll
  for (int j = KP - 1; j >= 0; j--) {
    const float lji = Lrow[g][j][i];  // L[j][i]
    if (i == j) xs[j & 1][g] = b / lji;
    __syncthreads();
    const float xj = xs[j & 1][g];
    if (i < j) b -= lji * xj;
    if (i == j) b = xj;
  }
  if (active) a.U[(size_t)row * KP + i] = b;
}

template <int KP, int NO>
static void launch_plain_t(const SampleArgs& a, hipStream_t s) {
  constexpr int WG = KP < 64 ? 64 : KP;
  constexpr int G = WG / KP;
  const int grid = (a.rows.N + G - 1) / G;
  if (grid == 0) return;
  if (a.univariate)
    hipLaunchKernelGGL((k_sample_plain<KP, NO, true>), dim3(grid), dim3(WG), 0, s, a);
  else
    hipLaunchKernelGGL((k_sample_plain<KP, NO, false>), dim3(grid), dim3(WG), 0, s, a);
}

template <int KP>
static void launch_plain_k(const SampleArgs& a, hipStream_t s) {
  switch (a.rows.nother) {
    case 1: return launch_plain_t<KP, 1>(a, s);
    case 2: return launch_plain_t<KP, 2>(a, s);
    case 3: return launch_plain_t<KP, 3>(a, s);
    case 4: return launch_plain_t<KP, 4>(a, s);
    default:
      if (a.rows.nother <= kMaxOther) return launch_plain_t<KP, kMaxOther>(a, s);
      break;
  }
}

void launch_sample_plain(const SampleArgs& a, hipStream_t s) {
  switch (a.KP) {
    case 8: return launch_plain_k<8>(a, s);
    case 16: return launch_plain_k<16>(a, s);
    case 32: return launch_plain_k<32>(a, s);
    case 64: return launch_plain_k<64>(a, s);
    case 128: return launch_plain_k<128>(a, s);
    default: break;
  }
}

}  // namespace mgpu
