// linop.hip -- side-information link matrix: SpMM with K right-hand sides and
// the block conjugate gradient of solve_blockcg (linop.h:228-343), FP64.
//
// Vectors of the CG are stored feature-major: column f holds the nr
// right-hand-side entries of feature f contiguously (the reference's
// row-major [nrhs x nfeat] Eigen matrices, linop.h:249-266).  The SpMM
// F P^T (A_mul_Bx<N>, linop.h:374-403) gathers one such column per nonzero
// with NP lanes, 64/NP nonzeros per wave-step: an HBM/L2 gather-bound loop.
#include "common.h"
#include "kernels.h"

namespace mgpu {

constexpr int kSpmmThreads = 256;

template <int NP>
__global__ void __launch_bounds__(kSpmmThreads)
    k_spmm(CsrDev A, const double* __restrict__ in, int nr, double* __restrict__ out, const double* __restrict__ reg,
           const double* __restrict__ add) {
  constexpr int S = 64 / NP;
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * kSpmmThreads + threadIdx.x) >> 6;
  const int nw = (gridDim.x * kSpmmThreads) >> 6;
  const int comp = lane % NP, sub = lane / NP;
  for (int r = w; r < A.nrow; r += nw) {
    double acc = 0.0;
    const int beg = A.row_ptr[r], end = A.row_ptr[r + 1];
    if (comp < nr) {
      for (int e = beg + sub; e < end; e += S) {
        const double v = A.vals ? A.vals[e] : 1.0;
        acc += v * in[(size_t)A.cols[e] * nr + comp];
      }
    }
#pragma unroll
    for (int off = NP; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
    if (sub == 0 && comp < nr) {
      if (add) acc += (*reg) * add[(size_t)r * nr + comp];
      out[(size_t)r * nr + comp] = acc;
    }
  }
}

void launch_spmm(const CsrDev& A, const double* in, int nr, double* out, const double* reg, const double* add,
                 hipStream_t s) {
  if (A.nrow == 0) return;
  int grid = (A.nrow + 3) / 4;
  if (grid > 4096) grid = 4096;
  if (nr <= 8) hipLaunchKernelGGL((k_spmm<8>), dim3(grid), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add);
  else if (nr <= 16) hipLaunchKernelGGL((k_spmm<16>), dim3(grid), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add);
  else if (nr <= 32) hipLaunchKernelGGL((k_spmm<32>), dim3(grid), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add);
  else hipLaunchKernelGGL((k_spmm<64>), dim3(grid), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add);
}

// ------------------------------------------------------------------
// partial lower dot products  sum_f a[:,f] b[:,f]^T  (A_mul_Bt_omp_sym,
// linop.cpp:320-372), FP64, one partial per block
// ------------------------------------------------------------------
constexpr int kDotThreads = 256;
constexpr int kMaxNr = 40;

int dots_blocks(int n) {
  int b = (n + 1023) / 1024;
  if (b < 1) b = 1;
  if (b > 256) b = 256;
  return b;
}

__global__ void __launch_bounds__(kDotThreads)
    k_dots(const double* __restrict__ a, const double* __restrict__ b, int nr, int n, double* __restrict__ partial,
           const int* done) {
  if (done && *done) return;
  constexpr int ENT = (kMaxNr * kMaxNr + kDotThreads - 1) / kDotThreads;
  constexpr int TC = 32;
  __shared__ double ta[TC][kMaxNr + 1];
  __shared__ double tb[TC][kMaxNr + 1];
  const int t = threadIdx.x;
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  const int c0 = blockIdx.x * chunk, c1 = min(n, c0 + chunk);
  const int NN = nr * nr;
  int ii[ENT], jj[ENT];
  double acc[ENT];
#pragma unroll
  for (int s = 0; s < ENT; s++) {
    const int q = t + s * kDotThreads;
    ii[s] = q < NN ? q % nr : 0;
    jj[s] = q < NN ? q / nr : 0;
    acc[s] = 0.0;
  }
  for (int base = c0; base < c1; base += TC) {
    const int nc = min(TC, c1 - base);
    for (int q = t; q < nc * nr; q += kDotThreads) {
      const int c = q / nr, i = q % nr;
      ta[c][i] = a[(size_t)(base + c) * nr + i];
      tb[c][i] = b[(size_t)(base + c) * nr + i];
    }
    __syncthreads();
    for (int c = 0; c < nc; c++) {
#pragma unroll
      for (int s = 0; s < ENT; s++) acc[s] += ta[c][ii[s]] * tb[c][jj[s]];
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < ENT; s++) {
    const int q = t + s * kDotThreads;
    if (q < NN) partial[(size_t)blockIdx.x * NN + q] = acc[s];
  }
}

void launch_dots(const double* a, const double* b, int nr, int n, double* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_dots, dim3(nblk), dim3(kDotThreads), 0, s, a, b, nr, n, partial, (const int*)nullptr);
}

// ------------------------------------------------------------------
// small nr x nr steps (one wave)
// small layout: RtR | RtR2 | PtKP | A | Psi (nr*nr each) | norms (nr) |
//               flags: done, iters (as doubles)
// ------------------------------------------------------------------
struct SmallView {
  double *RtR, *RtR2, *PtKP, *A, *Psi, *norms, *flags;
  __device__ SmallView(double* s, int nr) {
    const int NN = nr * nr;
    RtR = s;
    RtR2 = s + NN;
    PtKP = s + 2 * NN;
    A = s + 3 * NN;
    Psi = s + 4 * NN;
    norms = s + 5 * NN;
    flags = norms + nr;
  }
};

// reduce partials -> lower-symmetric matrix in LDS
__device__ void reduce_sym(const double* partial, int nblk, int nr, double* M) {
  const int NN = nr * nr;
  for (int q = threadIdx.x; q < NN; q += blockDim.x) {
    const int i = q % nr, j = q / nr;
    const int src = (i >= j) ? q : (j + i * nr);  // lower triangle (makeSymmetric)
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[(size_t)b * NN + src];
    M[q] = s;
  }
  __syncthreads();
}

// X = (M^{-1} Q)^T with M factored as Eigen's unblocked LLT does it
// (left-looking; stops at the first non-positive pivot and solves with the
// partially factored lower triangle).  The reference's block CG relies on
// that behaviour once the block Krylov space is exhausted.  All nr x nr
// operands live in LDS.
__device__ void small_solve_t(const double* M, const double* Q, double* X, int nr, double* L, int* sstop) {
  const int t = threadIdx.x;
  for (int q = t; q < nr * nr; q += blockDim.x) L[q] = M[q];
  if (t == 0) *sstop = 0;
  __syncthreads();
  for (int k = 0; k < nr; k++) {
    if (t == 0) {
      double x = L[k + k * nr];
      for (int j = 0; j < k; j++) x -= L[k + j * nr] * L[k + j * nr];
      if (x <= 0.0) *sstop = 1;
      else L[k + k * nr] = sqrt(x);
    }
    __syncthreads();
    if (*sstop) break;
    const double d = L[k + k * nr];
    for (int i = k + 1 + t; i < nr; i += blockDim.x) {
      double s = L[i + k * nr];
      for (int j = 0; j < k; j++) s -= L[i + j * nr] * L[k + j * nr];
      L[i + k * nr] = s / d;
    }
    __syncthreads();
  }
  // column j of S = M^{-1} Q; X[j][i] = S[i][j]
  for (int j = t; j < nr; j += blockDim.x) {
    double x[kMaxNr];
    for (int i = 0; i < nr; i++) {
      double s = Q[i + j * nr];
      for (int k = 0; k < i; k++) s -= L[i + k * nr] * x[k];
      x[i] = s / L[i + i * nr];
    }
    for (int i = nr - 1; i >= 0; i--) {
      double s = x[i];
      for (int k = i + 1; k < nr; k++) s -= L[k + i * nr] * x[k];
      x[i] = s / L[i + i * nr];
    }
    for (int i = 0; i < nr; i++) X[j + i * nr] = x[i];
  }
  __syncthreads();
}

// norms from diag(B B^T): small.norms; flags reset
__global__ void k_cg_norms(const double* partial, int nblk, int nr, double* small) {
  SmallView v(small, nr);
  const int NN = nr * nr;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[(size_t)b * NN + i + i * nr];
    v.norms[i] = sqrt(s);
  }
  if (threadIdx.x == 0) {
    v.flags[0] = 0.0;  // done
    v.flags[1] = 0.0;  // iterations
    v.flags[2] = 0.0;  // failure
  }
}

__global__ void k_cg_setup(const double* __restrict__ B, int nr, int n, const double* small, double* X, double* R,
                           double* P) {
  SmallView v(const_cast<double*>(small), nr);
  const size_t tot = (size_t)nr * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nr);
    const double r = B[q] * (1.0 / v.norms[i]);
    X[q] = 0.0;
    R[q] = r;
    P[q] = r;
  }
}

__global__ void k_cg_rtr(const double* partial, int nblk, int nr, double* small) {
  SmallView v(small, nr);
  reduce_sym(partial, nblk, nr, v.RtR);
}

__global__ void __launch_bounds__(64) k_cg_step_a(const double* partial, int nblk, int nr, double* small) {
  SmallView v(small, nr);
  if (v.flags[0] != 0.0) return;
  __shared__ double M[kMaxNr * kMaxNr], Q[kMaxNr * kMaxNr], L[kMaxNr * kMaxNr], Xs[kMaxNr * kMaxNr];
  __shared__ int sstop;
  reduce_sym(partial, nblk, nr, M);
  for (int q = threadIdx.x; q < nr * nr; q += blockDim.x) {
    v.PtKP[q] = M[q];
    Q[q] = v.RtR[q];
  }
  __syncthreads();
  small_solve_t(M, Q, Xs, nr, L, &sstop);
  for (int q = threadIdx.x; q < nr * nr; q += blockDim.x) v.A[q] = Xs[q];
}

template <int MODE>  // 0: X += A P, R -= A KP;  1: P = R + Psi P
__global__ void __launch_bounds__(256)
    k_cg_update(int nr, int n, const double* small, double* X, double* R, double* P, const double* KPm) {
  SmallView v(const_cast<double*>(small), nr);
  if (v.flags[0] != 0.0) return;  // converged or failed earlier
  __shared__ double Am[kMaxNr * kMaxNr];
  __shared__ double cp[8][kMaxNr], ck[8][kMaxNr];
  const double* src = MODE == 0 ? v.A : v.Psi;
  for (int q = threadIdx.x; q < nr * nr; q += blockDim.x) Am[q] = src[q];
  __syncthreads();
  const int cl = threadIdx.x / 32, i = threadIdx.x % 32;  // 8 columns x 32 lanes
  for (int base = blockIdx.x * 8; base < n; base += gridDim.x * 8) {
    const int f = base + cl;
    if (f < n) {
      for (int q = i; q < nr; q += 32) {
        cp[cl][q] = P[(size_t)f * nr + q];
        if (MODE == 0) ck[cl][q] = KPm[(size_t)f * nr + q];
      }
    }
    __syncthreads();
    if (f < n) {
      for (int r = i; r < nr; r += 32) {
        double sx = 0.0, sr = 0.0;
        for (int k = 0; k < nr; k++) {
          sx += Am[r + k * nr] * cp[cl][k];
          if (MODE == 0) sr += Am[r + k * nr] * ck[cl][k];
        }
        const size_t q = (size_t)f * nr + r;
        if (MODE == 0) {
          X[q] += sx;
          R[q] -= sr;
        } else {
          P[q] = R[q] + sx;
        }
      }
    }
    __syncthreads();
  }
}

// convergence test on diag(R R^T) and Psi = (RtR^{-1} RtR2)^T
__global__ void __launch_bounds__(64) k_cg_step_b(const double* partial, int nblk, int nr, double* small,
                                                   double tolsq) {
  SmallView v(small, nr);
  if (v.flags[0] != 0.0) return;
  __shared__ double M[kMaxNr * kMaxNr], Q[kMaxNr * kMaxNr], L[kMaxNr * kMaxNr], Xs[kMaxNr * kMaxNr];
  __shared__ int sstop, conv;
  if (threadIdx.x == 0) conv = 1;
  reduce_sym(partial, nblk, nr, Q);  // RtR2
  for (int i = threadIdx.x; i < nr; i += blockDim.x)
    if (!(Q[i + i * nr] < tolsq)) atomicAnd(&conv, 0);
  __syncthreads();
  if (conv) {
    if (threadIdx.x == 0) v.flags[0] = 1.0;
    return;
  }
  for (int q = threadIdx.x; q < nr * nr; q += blockDim.x) M[q] = v.RtR[q];
  __syncthreads();
  small_solve_t(M, Q, Xs, nr, L, &sstop);
  for (int q = threadIdx.x; q < nr * nr; q += blockDim.x) {
    v.Psi[q] = Xs[q];
    v.RtR[q] = Q[q];  // R R' = R2 R2' (linop.h:329)
  }
  if (threadIdx.x == 0) v.flags[1] += 1.0;
}

__global__ void k_cg_finish(int nr, int n, const double* small, double* X) {
  SmallView v(const_cast<double*>(small), nr);
  const size_t tot = (size_t)nr * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x)
    X[q] *= v.norms[q % nr];
}

static int grid_for(size_t work, int per) {
  size_t g = (work + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

void launch_cg_init(const CgState& c, const double* B, hipStream_t s) {
  launch_dots(B, B, c.nr, c.n, c.partial, c.nblk, s);
  hipLaunchKernelGGL(k_cg_norms, dim3(1), dim3(64), 0, s, c.partial, c.nblk, c.nr, c.small);
  hipLaunchKernelGGL(k_cg_setup, dim3(grid_for((size_t)c.nr * c.n, 256)), dim3(256), 0, s, B, c.nr, c.n,
                     (const double*)c.small, c.X, c.R, c.P);
  launch_dots(c.R, c.R, c.nr, c.n, c.partial, c.nblk, s);
  hipLaunchKernelGGL(k_cg_rtr, dim3(1), dim3(64), 0, s, c.partial, c.nblk, c.nr, c.small);
}

void launch_cg_step_a(const CgState& c, hipStream_t s) {
  hipLaunchKernelGGL(k_dots, dim3(c.nblk), dim3(kDotThreads), 0, s, (const double*)c.P, (const double*)c.KPm, c.nr,
                     c.n, c.partial, (const int*)nullptr);
  hipLaunchKernelGGL(k_cg_step_a, dim3(1), dim3(64), 0, s, c.partial, c.nblk, c.nr, c.small);
}

void launch_cg_update_xr(const CgState& c, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_update<0>), dim3(grid_for(c.n, 8)), dim3(256), 0, s, c.nr, c.n, (const double*)c.small,
                     c.X, c.R, c.P, (const double*)c.KPm);
}

void launch_cg_step_b(const CgState& c, double tolsq, hipStream_t s) {
  launch_dots(c.R, c.R, c.nr, c.n, c.partial, c.nblk, s);
  hipLaunchKernelGGL(k_cg_step_b, dim3(1), dim3(64), 0, s, c.partial, c.nblk, c.nr, c.small, tolsq);
}

void launch_cg_update_p(const CgState& c, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_update<1>), dim3(grid_for(c.n, 8)), dim3(256), 0, s, c.nr, c.n, (const double*)c.small,
                     c.X, c.R, c.P, (const double*)c.KPm);
}

void launch_cg_finish(const CgState& c, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_finish, dim3(grid_for((size_t)c.nr * c.n, 256)), dim3(256), 0, s, c.nr, c.n,
                     (const double*)c.small, c.X);
}

// ------------------------------------------------------------------
// sample_beta (latentprior.cpp:272-293) draws, wave per column:
//   tmp[:,n]  = U[:,n] + chol(Lambda)^{-T} z_n - mu
//   Fty[:,f] += sqrt(lambda_beta) chol(Lambda)^{-T} z'_f
// MvNormal_prec_omp (mvnormal.cpp:186-197).
// ------------------------------------------------------------------
__device__ void wave_back_solve(const double* __restrict__ L, int K, double x[2], int lane) {
  // x holds components lane and lane+64; solves L^T x = z in place
  for (int j = K - 1; j >= 0; j--) {
    const int owner = j & 63, slot = j >> 6;
    double xj = __shfl(slot ? x[1] : x[0], owner, 64);
    xj = xj / L[j + j * K];
    if (lane == owner) x[slot] = xj;
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      if (i < j) x[s] -= L[j + i * K] * xj;
    }
  }
}

template <int WHICH>  // 0: U columns, 1: feature columns
__global__ void __launch_bounds__(256) k_beta_draw(BetaArgs a, double* out) {
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int nw = (gridDim.x * 256) >> 6;
  const int it = *a.iter;
  const int ncol = WHICH == 0 ? a.N : a.nfeat;
  const uint32_t tg = make_tag(WHICH == 0 ? TAG_BETA_U : TAG_BETA_F, a.mode);
  const double sl = WHICH == 1 ? sqrt(*a.lambda_beta) : 0.0;
  for (int c = w; c < ncol; c += nw) {
    double x[2];
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      x[s] = i < a.K ? normal_f64(a.seed, (uint32_t)c, (uint32_t)i, (uint32_t)it, tg) : 0.0;
    }
    wave_back_solve(a.Lchol, a.K, x, lane);
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      if (i < a.K) {
        const size_t q = (size_t)c * a.K + i;
        if (WHICH == 0) out[q] = (double)a.U[(size_t)c * a.KP + i] + x[s] - a.mu[i];
        else out[q] += sl * x[s];
      }
    }
  }
}

void launch_beta_tmp(const BetaArgs& a, double* tmp, hipStream_t s) {
  hipLaunchKernelGGL((k_beta_draw<0>), dim3(grid_for(a.N, 4)), dim3(256), 0, s, a, tmp);
}
void launch_beta_noise_add(const BetaArgs& a, double* Fty, hipStream_t s) {
  hipLaunchKernelGGL((k_beta_draw<1>), dim3(grid_for(a.nfeat, 4)), dim3(256), 0, s, a, Fty);
}

// Uhat[:,n] = sum_f F[n,f] beta[:,f] (compute_uhat, linop.h:196-198) and the
// per-row prior term Lambda (mu + Uhat_n) used by MacauPrior::sample_latents
// (latentprior.cpp:218-227), wave per row.
__global__ void __launch_bounds__(256)
    k_uhat_finish(CsrDev F, const double* __restrict__ beta, int K, int KP, const double* __restrict__ mu,
                  const double* __restrict__ Lambda, float* __restrict__ Uhat, float* __restrict__ Lmu_rows) {
  __shared__ double us[4][128];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int w = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int nw = (gridDim.x * 256) >> 6;
  for (int r = w; r < F.nrow; r += nw) {
    double acc[2] = {0.0, 0.0};
    const int beg = F.row_ptr[r], end = F.row_ptr[r + 1];
    for (int e = beg; e < end; e++) {
      const double v = F.vals ? F.vals[e] : 1.0;
      const double* bcol = beta + (size_t)F.cols[e] * K;
      for (int s = 0; s < 2; s++) {
        const int i = lane + 64 * s;
        if (i < K) acc[s] += v * bcol[i];
      }
    }
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      if (i < K) us[wl][i] = mu[i] + acc[s];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      if (i < KP) {
        float uh = 0.f, lm = 0.f;
        if (i < K) {
          double t = 0.0;
          for (int j = 0; j < K; j++) t += Lambda[i + j * K] * us[wl][j];
          uh = (float)acc[s];
          lm = (float)t;
        }
        Uhat[(size_t)r * KP + i] = uh;
        Lmu_rows[(size_t)r * KP + i] = lm;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

void launch_uhat_finish(const CsrDev& F, const double* beta, int K, int KP, const double* mu, const double* Lambda,
                        float* Uhat, float* Lmu_rows, hipStream_t s) {
  if (F.nrow == 0) return;
  hipLaunchKernelGGL(k_uhat_finish, dim3(grid_for(F.nrow, 4)), dim3(256), 0, s, F, beta, K, KP, mu, Lambda, Uhat,
                     Lmu_rows);
}

// lambda_beta ~ Gamma(b, c) with posterior_lambda_beta (latentprior.cpp:317-331),
// nu0 = 1e-3, mu0 = 1 (latentprior.cpp:212-215)
__global__ void k_lambda_beta(int K, double nfeat, const double* bbt_partial, int nblk, const double* Lambda,
                              double* lambda_beta, uint64_t seed, const int* iter, int mode) {
  __shared__ double red[256];
  const int KK = K * K;
  double tr = 0.0;
  for (int q = threadIdx.x; q < KK; q += 256) {
    const int i = q % K, j = q / K;
    const int src = (i >= j) ? q : (j + i * K);
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += bbt_partial[(size_t)b * (K + KK) + K + src];
    tr += s * Lambda[j + i * K];
  }
  red[threadIdx.x] = tr;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nu = 1e-3, mu = 1.0;
    const double nux = nu + (double)K * nfeat;
    const double mux = mu * nux / (nu + mu * red[0]);
    const double b = nux / 2.0, c = 2.0 * mux / nux;
    *lambda_beta = gamma_f64(seed, b, 0u, (uint32_t)*iter, make_tag(TAG_LAMBDA_BETA, mode)) * c;
  }
}

void launch_lambda_beta(int K, double nfeat, const double* bbt_partial, int nblk, const double* Lambda,
                        double* lambda_beta, uint64_t seed, const int* iter, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_lambda_beta, dim3(1), dim3(256), 0, s, K, nfeat, bbt_partial, nblk, Lambda, lambda_beta,
                     seed, iter, mode);
}

}  // namespace mgpu

namespace mgpu {

__global__ void k_rows_copy(const double* __restrict__ src, int lds, double* __restrict__ dst, int ldd, int r0s,
                            int r0d, int nr, int n) {
  const size_t tot = (size_t)nr * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nr);
    const size_t c = q / nr;
    dst[c * ldd + r0d + i] = src[c * lds + r0s + i];
  }
}

void launch_rows_copy(const double* src, int lds, double* dst, int ldd, int r0s, int r0d, int nr, int n,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_rows_copy, dim3(grid_for((size_t)nr * n, 256)), dim3(256), 0, s, src, lds, dst, ldd, r0s, r0d,
                     nr, n);
}

template <typename T>
__global__ void k_sumsq(const T* __restrict__ x, size_t n, double* partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    const double v = (double)x[q];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

void launch_sumsq_f32(const float* x, size_t n, double* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_sumsq<float>), dim3(nblk), dim3(256), 0, s, x, n, partial);
}
void launch_sumsq_f64(const double* x, size_t n, double* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_sumsq<double>), dim3(nblk), dim3(256), 0, s, x, n, partial);
}

}  // namespace mgpu
