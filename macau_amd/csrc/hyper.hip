// hyper.hip -- FP64 hyperprior draw, adaptive noise, evaluation, RNG probes.
//
// Normal-Wishart conjugate update (CondNormalWishart, mvnormal.cpp:314-336):
//   mu_c = (kappa mu0 + sum U) / (kappa + N),  kappa_c = kappa + N,
//   nu_c = nu + N,  X = T + U U^T + kappa mu0 mu0^T - kappa_c mu_c mu_c^T,
//   T_c = X^{-1};  Lambda ~ Wishart(T_c, nu_c) by Bartlett
//   (mvnormal.cpp:217-266);  mu ~ N(mu_c, (kappa_c Lambda)^{-1})
//   (MvNormal_prec, mvnormal.cpp:169-184).
// The sufficient statistics are accumulated over columns by many work-groups
// in FP64 and reduced in place into the first partial row; the K x K draw
// then runs in one work-group per mode.
#include "common.h"
#include "kernels.h"

namespace mgpu {

constexpr int kMaxStatBlocks = 512;

int hyper_stats_blocks(int N) {
  int b = (N + 127) / 128;
  if (b < 1) b = 1;
  if (b > kMaxStatBlocks) b = kMaxStatBlocks;
  return b;
}

template <int KP>
struct StatShape {
  static constexpr int TD = KP / 4;                       // 4 x 4 tiles per dimension
  static constexpr int GT = TD * TD < 64 ? 64 : TD * TD;  // threads per column group
  static constexpr int G = GT >= 256 ? 1 : 256 / GT;      // column groups per block
  static constexpr int THREADS = GT * G;
  static constexpr int TC = 16;                           // columns per LDS tile
  static constexpr int EPT = TC * KP / GT;                // tile elements loaded per thread
  static constexpr int TILE = G * 2 * TC * KP;            // doubles, double buffered
  static constexpr int RED = G > 1 ? G * (KP + KP * KP) : 0;
  static constexpr int SMEM = TILE > RED ? TILE : RED;
};

// Sufficient statistics of x_n = U[:, n] - sub[:, n] (sub optional, FP32,
// same ld) over the block's run of columns, in FP64:
//   partial[blk] = (sum_n x_n (K), sum_n x_n x_n^T (K x K, col-major)).
// The block's columns are split between G column groups; each group streams
// its columns through a double-buffered LDS tile of 16 columns (the loads of
// tile t + 1 are in flight while tile t is used), and thread (ti, tj) of a
// group owns rows 4ti..4ti+3 x columns 4tj..4tj+3 of the K x K sum in
// registers.  The groups' sums are combined in group order.
template <typename T, int KP>
__global__ void __launch_bounds__(StatShape<KP>::THREADS)
    k_hyper_stats(const T* __restrict__ U, const float* __restrict__ sub, int ld, int N, int K, int cols,
                  double* __restrict__ partial) {
  using S = StatShape<KP>;
  constexpr int TD = S::TD, GT = S::GT, G = S::G, TC = S::TC, EPT = S::EPT;
  __shared__ double smem[S::SMEM];
  const int t = threadIdx.x, g = t / GT, gt = t % GT;
  const int ti = gt % TD, tj = gt / TD;
  const bool act = gt < TD * TD;
  const int b0 = blockIdx.x * cols, b1 = min(N, b0 + cols);
  const int per = (cols + G - 1) / G;
  const int c0 = min(b1, b0 + g * per), c1 = min(b1, c0 + per);
  const int ntiles = (per + TC - 1) / TC;
  double* tile = smem + (size_t)g * 2 * TC * KP;
  double reg[EPT];
  auto load = [&](int tt) {
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int f = gt + e * GT, c = f / KP, i = f % KP;
      const int col = c0 + tt * TC + c;
      double x = 0.0;
      if (col < c1 && i < K) {
        x = (double)U[(size_t)col * ld + i];
        if (sub) x -= (double)sub[(size_t)col * ld + i];
      }
      reg[e] = x;
    }
  };
  double acc[4][4], sum[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    sum[r] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) acc[r][q] = 0.0;
  }
  load(0);
  for (int tt = 0; tt < ntiles; tt++) {
    double* tb = tile + (tt & 1) * TC * KP;
#pragma unroll
    for (int e = 0; e < EPT; e++) tb[gt + e * GT] = reg[e];
    __syncthreads();
    if (tt + 1 < ntiles) load(tt + 1);
    if (act) {
#pragma unroll 4
      for (int c = 0; c < TC; c++) {
        const double* x = tb + c * KP;
        double xi[4], xj[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          xi[r] = x[4 * ti + r];
          xj[r] = x[4 * tj + r];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
          sum[r] += xi[r];
#pragma unroll
          for (int q = 0; q < 4; q++) acc[r][q] += xi[r] * xj[q];
        }
      }
    }
  }
  const int KK = K * K;
  double* out = partial + (size_t)blockIdx.x * (K + KK);
  if (G == 1) {
    if (!act) return;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 4 * ti + r;
      if (i >= K) continue;
      if (tj == 0) out[i] = sum[r];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = 4 * tj + q;
        if (j < K) out[K + i + j * K] = acc[r][q];
      }
    }
    return;
  }
  __syncthreads();  // tiles are dead: reuse the LDS for the group sums
  double* red = smem + (size_t)g * (K + KK);
  if (act) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 4 * ti + r;
      if (i >= K) continue;
      if (tj == 0) red[i] = sum[r];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = 4 * tj + q;
        if (j < K) red[K + i + j * K] = acc[r][q];
      }
    }
  }
  __syncthreads();
  for (int q = t; q < K + KK; q += S::THREADS) {
    double v = 0.0;
#pragma unroll
    for (int h = 0; h < G; h++) v += smem[(size_t)h * (K + KK) + q];
    out[q] = v;
  }
}

// partial[0][q] = sum_b partial[b][q] in a fixed order: 16 interleaved
// strands of blocks (strand g takes b = g, g + 16, ..., eight loads in
// flight), combined in strand order.  A work-group owns 16 entries q, so
// each load instruction reads 16 consecutive entries of 4 blocks.
constexpr int kRedStrands = 16, kRedEntries = 256 / kRedStrands;

__global__ void __launch_bounds__(256) k_stats_reduce(double* partial, int nblk, int width) {
  __shared__ double red[kRedStrands][kRedEntries];
  const int e = threadIdx.x % kRedEntries, g = threadIdx.x / kRedEntries;
  const int q = blockIdx.x * kRedEntries + e;
  double s = 0.0;
  if (q < width) {
    int b = g;
    for (; b + 7 * kRedStrands < nblk; b += 8 * kRedStrands) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = partial[(size_t)(b + kRedStrands * u) * width + q];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; b < nblk; b += kRedStrands) s += partial[(size_t)b * width + q];
  }
  red[g][e] = s;
  __syncthreads();
  if (g == 0 && q < width) {
    double t = red[0][e];
#pragma unroll
    for (int h = 1; h < kRedStrands; h++) t += red[h][e];
    partial[q] = t;
  }
}

void launch_partial_reduce(double* partial, int nblk, int width, hipStream_t s) {
  if (nblk > 1)
    hipLaunchKernelGGL(k_stats_reduce, dim3((width + kRedEntries - 1) / kRedEntries), dim3(256), 0, s, partial, nblk,
                       width);
}

template <typename T>
static void launch_stats(const T* U, const float* sub, int ld, int N, int K, double* partial, int nblk,
                         hipStream_t s) {
  const int KP = pad_latent(K);
  const int cols = (N + nblk - 1) / nblk;
  switch (KP) {
#define MGPU_STATS_CASE(P)                                                                                   \
  case P:                                                                                                    \
    hipLaunchKernelGGL((k_hyper_stats<T, P>), dim3(nblk), dim3(StatShape<P>::THREADS), 0, s, U, sub, ld, N, K, \
                       cols, partial);                                                                       \
    break;
    MGPU_STATS_CASE(8)
    MGPU_STATS_CASE(16)
    MGPU_STATS_CASE(32)
    MGPU_STATS_CASE(64)
    MGPU_STATS_CASE(128)
#undef MGPU_STATS_CASE
    default: break;
  }
  launch_partial_reduce(partial, nblk, K + K * K, s);
}

void launch_hyper_stats_f32(const float* U, const float* sub, int ld, int N, int K, double* partial, int nblk,
                            hipStream_t s) {
  launch_stats<float>(U, sub, ld, N, K, partial, nblk, s);
}
void launch_hyper_stats_f64(const double* U, int ld, int N, int K, double* partial, int nblk, hipStream_t s) {
  launch_stats<double>(U, nullptr, ld, N, K, partial, nblk, s);
}

// ------------------------------------------------------------------
// Block-cooperative dense FP64 helpers (col-major, ld = K).
// ------------------------------------------------------------------
constexpr int kNWThreads = 256;

// A = L L^T in place (lower), right-looking, upper part zeroed.  Returns
// false if a pivot is not positive.
__device__ bool blk_chol(double* A, int K, int* sflag) {
  const int t = threadIdx.x;
  for (int k = 0; k < K; k++) {
    if (t == 0) {
      const double d = A[k + k * K];
      if (!(d > 0.0)) *sflag = 1;
      A[k + k * K] = sqrt(d > 0.0 ? d : 1.0);
    }
    __syncthreads();
    const double inv = 1.0 / A[k + k * K];
    for (int i = k + 1 + t; i < K; i += kNWThreads) A[i + k * K] *= inv;
    __syncthreads();
    const int m = K - k - 1;
    for (int q = t; q < m * m; q += kNWThreads) {
      const int i = k + 1 + q % m, j = k + 1 + q / m;
      if (j <= i) A[i + j * K] -= A[i + k * K] * A[j + k * K];
    }
    __syncthreads();
  }
  for (int q = t; q < K * K; q += kNWThreads) {
    const int i = q % K, j = q / K;
    if (i < j) A[q] = 0.0;
  }
  __syncthreads();
  return *sflag == 0;
}

// Normal-Wishart draw of one mode (CondNormalWishart, mvnormal.cpp:314-336).
// With X = T + S + kappa mu0 mu0^T - kappa_c mu_c mu_c^T and J the reversal
// permutation, M = chol(J X J) gives chol(X^{-1}) = J M^{-T} J (the lower
// Cholesky factor is unique), so the Bartlett draw
//   Lambda = L_T (C^T C) L_T^T,  L_T = chol(T_c) = chol(X^{-1})
// (Wishart, mvnormal.cpp:243-266) becomes Lambda = B B^T with
// B = L_T C^T = J M^{-T} J C^T.  B is lower triangular with a positive
// diagonal, i.e. B = chol(Lambda), hence chol(kappa_c Lambda) = sqrt(kappa_c) B
// and mu = mu_c + B^{-T} z / sqrt(kappa_c) (MvNormal_prec, mvnormal.cpp:169-184).
// These are the reference's quantities computed through one Cholesky and one
// triangular solve instead of an explicit inverse, three Cholesky factors and
// three products.
__device__ void nw_draw(const NWArgs& a, double* base) {
  __shared__ int sflag;
  const int K = a.K, KK = K * K, t = threadIdx.x;
  double* A = base;       // X, then M
  double* Bm = A + KK;    // C (Bartlett, upper), then Z, then B
  double* D = Bm + KK;    // Lambda
  double* v_mu = D + KK;
  double* v_z = v_mu + K;
  double* v_nu = v_z + K;
  if (t == 0) sflag = 0;
  const int it = *a.iter;
  const double N = a.N;
  const double kappa_c = a.kappa + N, nu_c = a.nu + N;
  // 1. reduce the partial statistics (fixed block order, loads batched)
  const size_t stride = (size_t)(K + KK);
  for (int q = t; q < K + KK; q += kNWThreads) {
    double s = 0.0;
    int b = 0;
    for (; b + 8 <= a.nblk; b += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = a.partial[(size_t)(b + u) * stride + q];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; b < a.nblk; b++) s += a.partial[(size_t)b * stride + q];
    if (q < K) v_nu[q] = s;
    else D[q - K] = s;  // NS
  }
  __syncthreads();
  // 2. mu_c and X (mvnormal.cpp:322-325); Macau adds lambda_beta beta beta^T
  //    to T (latentprior.cpp:258-260).  A holds J X J.
  for (int k = t; k < K; k += kNWThreads) v_mu[k] = (a.kappa * a.mu0[k] + v_nu[k]) / (a.kappa + N);
  const double lb = a.lambda_beta ? *a.lambda_beta : 0.0;
  __syncthreads();
  for (int q = t; q < KK; q += kNWThreads) {
    const int i = q % K, j = q / K;
    double tq = a.WI[q];
    if (a.BBt_partial) {
      double s = 0.0;
      for (int b = 0; b < a.bbt_nblk; b++) s += a.BBt_partial[(size_t)b * stride + K + q];
      tq += lb * s;
    }
    const double x = tq + D[q] + a.kappa * a.mu0[i] * a.mu0[j] - kappa_c * v_mu[i] * v_mu[j];
    A[(K - 1 - i) + (K - 1 - j) * K] = x;
  }
  // Bartlett factor C (upper): C_ii = sqrt(chi2(nu_c - i)), C_ij ~ N(0,1)
  // (WishartUnit, mvnormal.cpp:217-241)
  for (int q = t; q < KK; q += kNWThreads) {
    const int i = q % K, j = q / K;
    double c = 0.0;
    if (i == j) {
      c = sqrt(2.0 * gamma_f64(a.seed, 0.5 * (nu_c - i), (uint32_t)i, (uint32_t)it, make_tag(TAG_WISHART_CHI, a.mode)));
    } else if (j > i) {
      c = normal_f64(a.seed, (uint32_t)i, (uint32_t)j, (uint32_t)it, make_tag(TAG_WISHART_Z, a.mode));
    }
    Bm[q] = c;
  }
  for (int k = t; k < K; k += kNWThreads)
    v_z[k] = normal_f64(a.seed, 0u, (uint32_t)k, (uint32_t)it, make_tag(TAG_NW_MU, a.mode));
  __syncthreads();
  // 3. M = chol(J X J)
  bool ok = blk_chol(A, K, &sflag);
  // 4. Z = M^{-T} R, R = J C^T (R[i][c] = C[c][K-1-i]); rows of Z from the
  //    bottom; Z[i][c] = 0 for i + c > K - 1 (exact structural zeros)
  for (int i = K - 1; i >= 0; i--) {
    for (int c = t; c < K; c += kNWThreads) {
      if (i + c <= K - 1) {
        double s = Bm[c + (K - 1 - i) * K];  // C[c][K-1-i]
        for (int k = i + 1; k + c <= K - 1; k++) s -= A[k + i * K] * D[k + c * K];
        D[i + c * K] = s / A[i + i * K];
      } else {
        D[i + c * K] = 0.0;
      }
    }
    __syncthreads();
  }
  // B = J Z (lower triangular)
  for (int q = t; q < KK; q += kNWThreads) {
    const int i = q % K, c = q / K;
    Bm[q] = D[(K - 1 - i) + c * K];
  }
  __syncthreads();
  // 5. Lambda = B B^T
  for (int q = t; q < KK; q += kNWThreads) {
    const int i = q % K, j = q / K;
    const int m = i < j ? i : j;
    double s = 0.0;
    for (int k = 0; k <= m; k++) s += Bm[i + k * K] * Bm[j + k * K];
    D[q] = s;
  }
  __syncthreads();
  // 6. mu = mu_c + B^{-T} z / sqrt(kappa_c)  (B^T upper, column-oriented)
  for (int j = K - 1; j >= 0; j--) {
    if (t == 0) v_z[j] = v_z[j] / Bm[j + j * K];
    __syncthreads();
    const double xj = v_z[j];
    for (int i = t; i < j; i += kNWThreads) v_z[i] -= Bm[j + i * K] * xj;
    __syncthreads();
  }
  const double isk = 1.0 / sqrt(kappa_c);
  for (int k = t; k < K; k += kNWThreads) {
    const double m = v_mu[k] + v_z[k] * isk;
    a.mu[k] = m;
    v_mu[k] = m;
  }
  __syncthreads();
  // 7. outputs: FP64 state and the FP32 sampler parameters
  const int KP = a.KP;
  for (int q = t; q < KK; q += kNWThreads) {
    a.Lambda[q] = D[q];
    if (a.Lchol) a.Lchol[q] = Bm[q];
  }
  for (int q = t; q < KP * KP; q += kNWThreads) {
    const int i = q % KP, j = q / KP;
    float v = (i == j) ? 1.f : 0.f;
    if (i < K && j < K) v = (float)D[i + j * K];
    a.Lambda_f[q] = v;
  }
  for (int i = t; i < KP; i += kNWThreads) {
    double s = 0.0;
    if (i < K)
      for (int j = 0; j < K; j++) s += D[i + j * K] * v_mu[j];
    a.Lmu_f[i] = (float)s;
  }
  if (t == 0 && !ok && a.fail) atomicMax(a.fail, kFailNormalWishart);
}

// ------------------------------------------------------------------
// The same draw in one work-group of four waves for K <= 64, every loop
// rolled over K with the matrices in LDS.  The factorisation and solves are
// chains of dependent FP64 operations on one wave, so their dot products
// carry four partial sums, the draws of the Bartlett factor (which the
// factorisation does not need) run on the other three waves meanwhile, and
// Lambda = B B^T is split over all four.
//   A: J X J, then M = chol(J X J) in its lower triangle (left-looking, lane
//      i = row i), then B (rows);
//   C: the Bartlett factor C (rows); Z = M^{-T} R, R = J C^T, is solved in
//      place (lane c = column c of Z; R[i][c] = C[c][K-1-i] is read once, at
//      the step that writes Z[i][c] to the same place), which leaves
//      C[c][m] = Z[K-1-m][c] = B[m][c], i.e. C = B^T with B = J Z.
// ------------------------------------------------------------------
constexpr int kNWW = 64;         // largest K of this draw
constexpr int kNWLd = kNWW + 1;  // row stride (doubles): lanes on their own rows hit distinct banks
constexpr int kNWThreadsW = 256;

struct NWWaveLds {
  double A[kNWW][kNWLd];
  double C[kNWW][kNWLd];
  double muc[kNWW];     // mu_c, then mu
  double z[kNWW];       // normals of the mu draw
  double invd[kNWW];    // 1 / M[k][k]
  double lm[4][kNWW];   // Lambda mu, per wave
};

// Philox block (row i, block b) of the strictly upper Bartlett normals
// C[i][4b .. 4b+3]: item t of the blocks holding any j with i < j < K, in
// block-column order
__device__ __forceinline__ bool bartlett_block(int t, int K, int& i, int& b) {
  for (int q = 0; 4 * q < K; q++) {
    const int rows = min(4 * q + 3, K - 1);  // i = 0 .. rows-1 have a j > i in the block
    if (t < rows) {
      i = t;
      b = q;
      return true;
    }
    t -= rows;
  }
  return false;
}

// 1/sqrt(d), d > 0: the hardware estimate (v_rsq_f64) refined by two
// Newton-Raphson steps for f(y) = y^-2 - d, y <- y (3 - d y^2) / 2, each of
// which roughly squares the relative error
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// sum_{j in [j0, j1)} x[j] y[j] with four partial sums
__device__ __forceinline__ double dot4(const double* x, const double* y, int j0, int j1) {
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    p0 += x[j] * y[j];
    p1 += x[j + 1] * y[j + 1];
    p2 += x[j + 2] * y[j + 2];
    p3 += x[j + 3] * y[j + 3];
  }
  for (; j < j1; j++) p0 += x[j] * y[j];
  return (p0 + p1) + (p2 + p3);
}

// Lambda = B B^T (B rows in A, zero beyond column K) on all waves: lane r =
// row r, wave w takes columns j = w, w + 4, ..., four at a time.  KB bounds
// the row length.
template <int KB>
__device__ __forceinline__ void nw_lambda(const NWArgs& a, NWWaveLds& s, int wave, int lane) {
  const int K = a.K, KP = a.KP;
  const int r = lane;
  double lm = 0.0;
  for (int j0 = 4 * wave; j0 < KP; j0 += 16) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double* bj[4];
#pragma unroll
    for (int u = 0; u < 4; u++) bj[u] = s.A[min(j0 + u, kNWW - 1)];
#pragma unroll 2
    for (int k = 0; k < KB; k += 8) {
      double br[8];
#pragma unroll
      for (int t = 0; t < 8; t++) br[t] = s.A[r][k + t];
#pragma unroll
      for (int t = 0; t < 8; t++)
#pragma unroll
        for (int u = 0; u < 4; u++) acc[u] += br[t] * bj[u][k + t];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = j0 + u;
      if (j >= KP) break;
      const bool in = r < K && j < K;
      if (in) {
        lm += acc[u] * s.muc[j];
        a.Lambda[r + j * K] = acc[u];
      }
      if (r < KP) a.Lambda_f[r + j * KP] = in ? (float)acc[u] : (r == j ? 1.f : 0.f);
    }
  }
  s.lm[wave][r] = lm;
}

__device__ void nw_draw_w(const NWArgs& a, NWWaveLds& s) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int K = a.K;
  const bool act = lane < K;
  const int it = *a.iter;
  const double N = a.N;
  const double kappa_c = a.kappa + N, nu_c = a.nu + N;
  bool bad = false;
  if (wave == 0) {
    // 1. mu_c (mvnormal.cpp:322-323) and the normals of the mu draw
    if (act) {
      s.z[lane] = normal_f64(a.seed, 0u, (uint32_t)lane, (uint32_t)it, make_tag(TAG_NW_MU, a.mode));
      s.muc[lane] = (a.kappa * a.mu0[lane] + a.partial[lane]) / kappa_c;
    }
    lds_wait();
    // 2. A = J X J, X = T + S + kappa mu0 mu0^T - kappa_c mu_c mu_c^T
    //    (mvnormal.cpp:324-325; Macau adds lambda_beta beta beta^T to T,
    //    latentprior.cpp:258-260); lane = row, loads coalesced along it
    const double lb = a.lambda_beta ? *a.lambda_beta : 0.0;
    const double* bbt = a.BBt_partial ? a.BBt_partial + K : a.partial + K;
    const double lbb = a.BBt_partial ? lb : 0.0;
    const int xi = act ? K - 1 - lane : 0;
    const double m0i = a.mu0[xi], mci = s.muc[xi];
#pragma unroll 8
    for (int c = 0; c < K; c++) {
      const int xj = K - 1 - c;
      const int e = xi + xj * K;
      const double x = a.WI[e] + lbb * bbt[e] + a.partial[K + e] + a.kappa * m0i * a.mu0[xj] - kappa_c * mci * s.muc[xj];
      if (act) s.A[lane][c] = x;
    }
    lds_wait();
    // 3. M = chol(A), left-looking: s_i = A[i][k] - sum_{j<k} M[i][j] M[k][j]
    //    on lane i, the pivot from lane k
    for (int k = 0; k < K; k++) {
      const double sk = s.A[lane][k] - dot4(s.A[lane], s.A[k], 0, k);
      const double d = rdlane(sk, k);
      const bool badp = !(d > 0.0);
      bad = bad || badp;
      const double rs = rsqrt_nr(badp ? 1.0 : d);
      if (act && lane >= k) s.A[lane][k] = sk * rs;  // M[i][k]
      s.invd[k] = rs;                               // every lane stores the same value
      lds_wait();
    }
  } else {
    // 4. Bartlett factor (waves 1-3): C_ii = sqrt(chi2(nu_c - i)),
    //    C_ij ~ N(0,1) for i < j < K (WishartUnit, mvnormal.cpp:217-241),
    //    zero below the diagonal; normals C[i][4b..4b+3] from one Philox block
    const int t0 = threadIdx.x - 64;
    for (int i = wave; i < K; i += 3)
      if (lane < i) s.C[i][lane] = 0.0;
    for (int t = t0;; t += kNWThreadsW - 64) {
      int i, b;
      if (!bartlett_block(t, K, i, b)) break;
      double zz[4];
      normal_f64x4(a.seed, (uint32_t)i, (uint32_t)b, (uint32_t)it, make_tag(TAG_WISHART_Z, a.mode), zz);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = 4 * b + u;
        if (j > i && j < K) s.C[i][j] = zz[u];
      }
    }
    for (int i = t0; i < K; i += kNWThreadsW - 64)
      s.C[i][i] = sqrt(2.0 * gamma_f64(a.seed, 0.5 * (nu_c - i), (uint32_t)i, (uint32_t)it,
                                       make_tag(TAG_WISHART_CHI, a.mode)));
  }
  __syncthreads();
  if (wave == 0) {
    // 5. Z = M^{-T} R from the bottom row up, lane c = column c, in place in C
    for (int i = K - 1; i >= 0; i--) {
      double p0 = s.C[lane][K - 1 - i], p1 = 0.0, p2 = 0.0, p3 = 0.0;
      int k = i + 1;
      for (; k + 4 <= K; k += 4) {
        p0 -= s.A[k][i] * s.C[lane][K - 1 - k];
        p1 -= s.A[k + 1][i] * s.C[lane][K - 2 - k];
        p2 -= s.A[k + 2][i] * s.C[lane][K - 3 - k];
        p3 -= s.A[k + 3][i] * s.C[lane][K - 4 - k];
      }
      for (; k < K; k++) p0 -= s.A[k][i] * s.C[lane][K - 1 - k];
      if (act) s.C[lane][K - 1 - i] = ((p0 + p1) + (p2 + p3)) * s.invd[i];
    }
    lds_wait();
    // 6. B rows into A (B[i][c] = C[c][i]), zero beyond column K
    for (int i = 0; i < K; i++) s.A[i][lane] = act ? s.C[lane][i] : 0.0;
    // 7. mu = mu_c + B^{-T} z / sqrt(kappa_c) (MvNormal_prec, mvnormal.cpp:169-184):
    //    B^T x = z from the last row up; lane r holds x_r, B[j][r] = C[r][j]
    double acc = act ? s.z[lane] : 0.0;
    const double invb = act ? 1.0 / s.C[lane][lane] : 0.0;
    for (int j = K - 1; j >= 0; j--) {
      const double xj = rdlane(acc * invb, j);
      acc = (lane < j) ? acc - s.C[lane][j] * xj : (lane == j ? xj : acc);
    }
    if (act) {
      const double mu = s.muc[lane] + acc / sqrt(kappa_c);
      s.muc[lane] = mu;
      a.mu[lane] = mu;
    }
    lds_wait();
    if (a.Lchol && act)
      for (int r = 0; r < K; r++) a.Lchol[lane + r * K] = s.A[lane][r];  // B[lane][r]
  }
  __syncthreads();
  // 8. Lambda = B B^T and Lambda mu
  if (K <= 32) nw_lambda<32>(a, s, wave, lane);
  else nw_lambda<64>(a, s, wave, lane);
  __syncthreads();
  if (wave == 0 && lane < a.KP)
    a.Lmu_f[lane] = (float)((s.lm[0][lane] + s.lm[1][lane]) + (s.lm[2][lane] + s.lm[3][lane]));
  if (threadIdx.x == 0 && bad && a.fail) atomicMax(a.fail, kFailNormalWishart);
}

__global__ void __launch_bounds__(kNWThreadsW) k_nw_draw_w(NWBatch b) {
  __shared__ NWWaveLds s;
  nw_draw_w(b.m[blockIdx.x], s);
}

// The same draw for 64 < K <= 128 with every matrix in LDS (nw_draw keeps
// them in a global workspace there: its per-step barriers then wait on
// global memory, ~2.5 ms at K = 128).  Triangles are packed: M = chol(J X J)
// lower, row i at i (i + 1) / 2; the Bartlett factor C upper, column j at
// j (j + 1) / 2.  Row i of Z = M^{-T} J C^T only needs column K - 1 - i of C
// and is written over it, and B = J Z is then the packed lower triangle
// read row by row; Lambda = B B^T takes M's place.  Every value goes through
// the same operations in the same order as in nw_draw.  LDS:
// (K (K + 1) + 3 K) doubles, 132 KB at K = 128.
__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

__global__ void __launch_bounds__(kNWThreads) k_nw_draw_lds(NWBatch b) {
  extern __shared__ double dyn[];
  const NWArgs& a = b.m[blockIdx.x];
  __shared__ int sflag;
  const int K = a.K, KK = K * K, t = threadIdx.x, T2 = tri(K);
  double* Mp = dyn;         // J X J, then M (packed lower), then Lambda (packed lower)
  double* Cp = Mp + T2;     // C (packed upper, by columns), then Z rows, i.e. B (packed lower)
  double* v_mu = Cp + T2;
  double* v_z = v_mu + K;
  double* v_nu = v_z + K;
  if (t == 0) sflag = 0;
  const int it = *a.iter;
  const double N = a.N;
  const double kappa_c = a.kappa + N, nu_c = a.nu + N;
  const size_t stride = (size_t)(K + KK);
  auto reduce = [&](int q) {  // the statistic q summed over the blocks, in block order
    double sum = 0.0;
    int bb = 0;
    for (; bb + 8 <= a.nblk; bb += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = a.partial[(size_t)(bb + u) * stride + q];
#pragma unroll
      for (int u = 0; u < 8; u++) sum += v[u];
    }
    for (; bb < a.nblk; bb++) sum += a.partial[(size_t)bb * stride + q];
    return sum;
  };
  // 1. sums of the columns, mu_c
  for (int q = t; q < K; q += kNWThreads) v_nu[q] = reduce(q);
  __syncthreads();
  for (int k = t; k < K; k += kNWThreads) v_mu[k] = (a.kappa * a.mu0[k] + v_nu[k]) / (a.kappa + N);
  const double lb = a.lambda_beta ? *a.lambda_beta : 0.0;
  __syncthreads();
  // 2. J X J (lower: entry (r, c) = X[K-1-r][K-1-c]), the Bartlett factor, z
  for (int e = t; e < T2; e += kNWThreads) {
    int r = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (tri(r) > e) r--;
    while (tri(r + 1) <= e) r++;
    const int c = e - tri(r);
    const int i = K - 1 - r, j = K - 1 - c, q = i + j * K;
    double tq = a.WI[q];
    if (a.BBt_partial) {
      double sb = 0.0;
      for (int bb = 0; bb < a.bbt_nblk; bb++) sb += a.BBt_partial[(size_t)bb * stride + K + q];
      tq += lb * sb;
    }
    Mp[e] = tq + reduce(K + q) + a.kappa * a.mu0[i] * a.mu0[j] - kappa_c * v_mu[i] * v_mu[j];
    // C[c'][j'] of packed slot e (column j' = r, row c' = c)
    double cv;
    if (c == r)
      cv = sqrt(2.0 * gamma_f64(a.seed, 0.5 * (nu_c - c), (uint32_t)c, (uint32_t)it, make_tag(TAG_WISHART_CHI, a.mode)));
    else
      cv = normal_f64(a.seed, (uint32_t)c, (uint32_t)r, (uint32_t)it, make_tag(TAG_WISHART_Z, a.mode));
    Cp[e] = cv;
  }
  for (int k = t; k < K; k += kNWThreads)
    v_z[k] = normal_f64(a.seed, 0u, (uint32_t)k, (uint32_t)it, make_tag(TAG_NW_MU, a.mode));
  __syncthreads();
  // 3. M = chol(J X J), right-looking in blocks of 4 columns, two barriers
  //    per block: every thread factors the 4 x 4 diagonal block itself
  //    (pivots <= 0 flagged and replaced by 1, as column by column), thread
  //    t solves row k0 + 4 + t of the block column, thread 0 stores the
  //    diagonal block's factor after the barrier (once no thread reads the
  //    old entries), and the rank-4 trailing update runs on a 16 x 16
  //    thread grid, each row's entries loaded before any is stored
  const int tx = t & 15, ty = t >> 4;
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int nb = K - k0 < 4 ? K - k0 : 4;
    double L4[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (c >= nb) break;
      double d = Mp[tri(k0 + c) + k0 + c];
#pragma unroll
      for (int m = 0; m < c; m++) d -= L4[c][m] * L4[c][m];
      if (t == 0 && !(d > 0.0)) sflag = 1;
      L4[c][c] = sqrt(d > 0.0 ? d : 1.0);
#pragma unroll
      for (int r = c + 1; r < 4; r++) {
        if (r >= nb) break;
        double v = Mp[tri(k0 + r) + k0 + c];
#pragma unroll
        for (int m = 0; m < c; m++) v -= L4[r][m] * L4[c][m];
        L4[r][c] = v / L4[c][c];
      }
    }
    for (int i = k0 + nb + t; i < K; i += kNWThreads) {
      double l[4];
#pragma unroll
      for (int c = 0; c < 4; c++) l[c] = c < nb ? Mp[tri(i) + k0 + c] : 0.0;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        if (c >= nb) break;
#pragma unroll
        for (int m = 0; m < c; m++) l[c] -= l[m] * L4[c][m];
        l[c] = l[c] / L4[c][c];
      }
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (c < nb) Mp[tri(i) + k0 + c] = l[c];
    }
    __syncthreads();
    if (t == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c <= r; c++)
          if (r < nb) Mp[tri(k0 + r) + k0 + c] = L4[r][c];
    }
    const int j0 = k0 + nb;
    for (int i = j0 + ty; i < K; i += 16) {
      double li[4];
#pragma unroll
      for (int c = 0; c < 4; c++) li[c] = c < nb ? Mp[tri(i) + k0 + c] : 0.0;
      double x[8], s4[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int j = j0 + tx + 16 * u;
        const bool in = j <= i;
        x[u] = in ? Mp[tri(i) + j] : 0.0;
        double sv = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) sv += (in && c < nb) ? li[c] * Mp[tri(j) + k0 + c] : 0.0;
        s4[u] = sv;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int j = j0 + tx + 16 * u;
        if (j <= i) Mp[tri(i) + j] = x[u] - s4[u];
      }
    }
    __syncthreads();
  }
  const bool ok = sflag == 0;
  // 4. Z = M^{-T} R, R = J C^T, rows from the bottom, right-looking in
  //    blocks of 4 rows: row i of R lives over column K - 1 - i of C (slots
  //    tri(K-1-i) + c, c <= K - 1 - i).  For a block of rows b0..b0+nb-1,
  //    thread c solves the block's 4 x 4 upper system for its column c
  //    (rows i with c <= K - 1 - i), then every row i' < b0 subtracts
  //    sum_k M_ki' Z_k over the block's rows k with c <= K - 1 - k (16 x 16
  //    thread grid, each row's entries loaded before any store)
  for (int b1 = K; b1 > 0; b1 -= 4) {
    const int b0 = b1 - 4 > 0 ? b1 - 4 : 0, nb = b1 - b0;
    for (int c = t; c <= K - 1 - b0; c += kNWThreads) {
      double z[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int q = 3; q >= 0; q--) {
        const int i = b0 + q;
        if (q >= nb || c > K - 1 - i) continue;
        double v = Cp[tri(K - 1 - i) + c];
#pragma unroll
        for (int q2 = q + 1; q2 < 4; q2++)
          if (q2 < nb && c <= K - 1 - (b0 + q2)) v -= Mp[tri(b0 + q2) + i] * z[q2];
        z[q] = v / Mp[tri(i) + i];
        Cp[tri(K - 1 - i) + c] = z[q];
      }
    }
    __syncthreads();
    const int cmax = K - 1 - b0;  // widest row of the block
    for (int i2 = ty; i2 < b0; i2 += 16) {
      double m[4];
#pragma unroll
      for (int q = 0; q < 4; q++) m[q] = q < nb ? Mp[tri(b0 + q) + i2] : 0.0;
      const int j2 = K - 1 - i2;
      double x[8], s4[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int c = tx + 16 * u;
        const bool in = c <= cmax;
        x[u] = in ? Cp[tri(j2) + c] : 0.0;
        double sv = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++)
          sv += (in && q < nb && c <= K - 1 - (b0 + q)) ? m[q] * Cp[tri(K - 1 - (b0 + q)) + c] : 0.0;
        s4[u] = sv;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int c = tx + 16 * u;
        if (c <= cmax) Cp[tri(j2) + c] = x[u] - s4[u];
      }
    }
    __syncthreads();
  }
  // B[i][c] = Z[K-1-i][c] = Cp[tri(i) + c] (c <= i).
  // 5. Lambda = B B^T (packed lower, over M), four partial sums per entry
  for (int e = t; e < T2; e += kNWThreads) {
    int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (tri(i) > e) i--;
    while (tri(i + 1) <= e) i++;
    const int j = e - tri(i);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = 0;
    for (; k + 4 <= j + 1; k += 4) {
      s0 += Cp[tri(i) + k] * Cp[tri(j) + k];
      s1 += Cp[tri(i) + k + 1] * Cp[tri(j) + k + 1];
      s2 += Cp[tri(i) + k + 2] * Cp[tri(j) + k + 2];
      s3 += Cp[tri(i) + k + 3] * Cp[tri(j) + k + 3];
    }
    for (; k <= j; k++) s0 += Cp[tri(i) + k] * Cp[tri(j) + k];
    Mp[e] = (s0 + s1) + (s2 + s3);
  }
  // 6. mu = mu_c + B^{-T} z / sqrt(kappa_c): back substitution on wave 0,
  //    lane l holding z_l and z_{l+64}, the solved value broadcast by a
  //    shuffle (no barrier per step)
  if (t < 64) {
    double z0 = v_z[t], z1 = t + 64 < K ? v_z[t + 64] : 0.0;
    for (int j = K - 1; j >= 0; j--) {
      const double zj = j >= 64 ? __shfl(z1, j - 64, 64) : __shfl(z0, j, 64);
      const double xj = zj / Cp[tri(j) + j];
      if (t == j) z0 = xj;
      if (t + 64 == j) z1 = xj;
      if (t < j) z0 -= Cp[tri(j) + t] * xj;
      if (t + 64 < j) z1 -= Cp[tri(j) + t + 64] * xj;
    }
    v_z[t] = z0;
    if (t + 64 < K) v_z[t + 64] = z1;
  }
  __syncthreads();
  const double isk = 1.0 / sqrt(kappa_c);
  for (int k = t; k < K; k += kNWThreads) {
    const double m = v_mu[k] + v_z[k] * isk;
    a.mu[k] = m;
    v_mu[k] = m;
  }
  __syncthreads();
  // 7. outputs
  const int KP = a.KP;
  for (int q = t; q < KK; q += kNWThreads) {
    const int i = q % K, j = q / K;
    a.Lambda[q] = i >= j ? Mp[tri(i) + j] : Mp[tri(j) + i];
    if (a.Lchol) a.Lchol[q] = i >= j ? Cp[tri(i) + j] : 0.0;
  }
  for (int q = t; q < KP * KP; q += kNWThreads) {
    const int i = q % KP, j = q / KP;
    float v = (i == j) ? 1.f : 0.f;
    if (i < K && j < K) v = (float)(i >= j ? Mp[tri(i) + j] : Mp[tri(j) + i]);
    a.Lambda_f[q] = v;
  }
  for (int i = t; i < KP; i += kNWThreads) {
    double sv = 0.0;
    if (i < K)
      for (int j = 0; j < K; j++) sv += (i >= j ? Mp[tri(i) + j] : Mp[tri(j) + i]) * v_mu[j];
    a.Lmu_f[i] = (float)sv;
  }
  if (t == 0 && !ok && a.fail) atomicMax(a.fail, kFailNormalWishart);
}

// one work-group per mode of the batch
__global__ void __launch_bounds__(kNWThreads) k_nw_draw(NWBatch b) {
  extern __shared__ double dyn[];
  const NWArgs& a = b.m[blockIdx.x];
  nw_draw(a, a.ws ? a.ws : dyn);
}

static size_t nw_lds_bytes(int K) { return (size_t)(3 * K * K + 3 * K) * sizeof(double); }
size_t nw_workspace_bytes(int K) { return nw_lds_bytes(K) <= 48 * 1024 ? 0 : nw_lds_bytes(K); }

void launch_nw_draw_batch(const NWBatch& b, hipStream_t s) {
  if (b.n <= 0) return;
  // one wave per mode when the statistics arrive reduced and K <= 64
  bool wave = true;
  for (int q = 0; q < b.n; q++) wave = wave && b.m[q].nblk == 1 && b.m[q].K <= kNWW;
  if (wave) {
    hipLaunchKernelGGL(k_nw_draw_w, dim3(b.n), dim3(kNWThreadsW), 0, s, b);
    return;
  }
  // 64 < K <= 128: the LDS-resident draw where the device grants a
  // work-group that much LDS (checked once per device)
  const int K = b.m[0].K;
  if (K > kNWW && K <= 128) {
    const size_t bytes = (size_t)(K * (K + 1) + 3 * K) * sizeof(double);
    static int granted[64] = {0};  // per device: 0 unknown, 1 yes, -1 no
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) {
      if (granted[dev] == 0) {
        int optin = 0;
        granted[dev] = -1;
        if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess &&
            (size_t)optin >= (size_t)(128 * 129 + 3 * 128) * sizeof(double) &&
            hipFuncSetAttribute((const void*)k_nw_draw_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)((size_t)(128 * 129 + 3 * 128) * sizeof(double))) == hipSuccess)
          granted[dev] = 1;
        (void)hipGetLastError();
      }
      if (granted[dev] == 1) {
        hipLaunchKernelGGL(k_nw_draw_lds, dim3(b.n), dim3(kNWThreads), bytes, s, b);
        return;
      }
    }
  }
  const bool lds = b.m[0].ws == nullptr;
  hipLaunchKernelGGL(k_nw_draw, dim3(b.n), dim3(kNWThreads), lds ? nw_lds_bytes(b.m[0].K) : 0, s, b);
}

void launch_nw_draw(const NWArgs& a, hipStream_t s) {
  NWBatch b;
  b.m[0] = a;
  b.n = 1;
  launch_nw_draw_batch(b, s);
}

// ------------------------------------------------------------------
// Adaptive noise: SSE over the training cells (noisemodels.cpp:55-124)
// ------------------------------------------------------------------
constexpr int kNoiseThreads = 256;

int noise_sse_blocks(int N0) {
  int b = (N0 + 3) / 4;
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return b;
}

template <int KP>
__global__ void __launch_bounds__(kNoiseThreads) k_noise_sse(NoiseArgs a) {
  __shared__ float us[kNoiseThreads / 64][KP];
  __shared__ double red[kNoiseThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = gridDim.x * (kNoiseThreads / 64);
  const int no = a.nmodes - 1;
  double acc = 0.0;
  for (int row = blockIdx.x * (kNoiseThreads / 64) + w; row < a.rows0.N; row += nw) {
    for (int i = lane; i < KP; i += 64) us[w][i] = a.U0[(size_t)row * KP + i];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int beg = a.rows0.row_ptr[row], end = a.rows0.row_ptr[row + 1];
    for (int e = beg + lane; e < end; e += 64) {
      float dot = 0.f;
      const int* ix = a.rows0.idx + (size_t)e * no;
      for (int d = 0; d < a.K; d++) {
        float v = us[w][d];
        for (int m = 0; m < no; m++) v *= a.others.f[m][(size_t)ix[m] * KP + d];
        dot += v;
      }
      const double r = (double)dot - (double)a.rows0.r[e];
      acc += r * r;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) red[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < kNoiseThreads / 64; q++) s += red[q];
    a.partial[blockIdx.x] = s;
  }
}

void launch_noise_sse(const NoiseArgs& a, hipStream_t s) {
  switch (a.KP) {
    case 8: hipLaunchKernelGGL((k_noise_sse<8>), dim3(a.nblk), dim3(kNoiseThreads), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_noise_sse<16>), dim3(a.nblk), dim3(kNoiseThreads), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_noise_sse<32>), dim3(a.nblk), dim3(kNoiseThreads), 0, s, a); break;
    case 64: hipLaunchKernelGGL((k_noise_sse<64>), dim3(a.nblk), dim3(kNoiseThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_noise_sse<128>), dim3(a.nblk), dim3(kNoiseThreads), 0, s, a); break;
  }
}

// alpha ~ Gamma(a0 + nnz/2, 1/(b0 + SSE/2)), clamped to alpha_max
__global__ void k_noise_draw(NoiseDrawArgs a) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < a.nblk; b += 256) s += a.partial[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double sse = red[0];
    const double aN = 0.5 + a.nnz / 2.0;
    const double bN = a.b0 + sse / 2.0;
    double al = gamma_f64(a.seed, aN, 0u, (uint32_t)*a.iter, make_tag(TAG_NOISE, 0)) / bN;
    if (al > a.alpha_max) al = a.alpha_max;
    *a.alpha = al;
    if (a.sse_out) *a.sse_out = sse;
  }
}

void launch_noise_draw(const NoiseDrawArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_noise_draw, dim3(1), dim3(256), 0, s, a);
}

// ------------------------------------------------------------------
// Evaluation: eval_rmse / eval_rmse_tensor (bpmfutils.h:86-114,
// bpmfutils.cpp:10-73) with Welford running mean and variance.
// A group of KP/4 lanes handles one test cell: each lane multiplies one
// float4 of the factor columns, the group sums by shuffles, and the group's
// first lane updates the FP64 running mean/variance and squared errors.
// ------------------------------------------------------------------
constexpr int kEvalThreads = 256;
int eval_blocks(int ntest) {
  int b = (ntest + kEvalThreads * 2 - 1) / (kEvalThreads * 2);
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return b;
}

template <int KP, int NM, bool PROBIT>
__global__ void __launch_bounds__(kEvalThreads) k_eval(EvalArgs a) {
  constexpr int G = KP / 4;  // lanes per test cell
  __shared__ double red[2][kEvalThreads / 64];
  const int it = *a.iter;
  const int n = it < a.burnin ? 0 : it - a.burnin;
  const int sub = threadIdx.x % G;
  const int cells_per_block = kEvalThreads / G;
  // U consecutive cells per lane group and loop step: their index loads,
  // then their factor-row loads, are issued together, and after the
  // lane-group sums lane u of the group finishes cell u, so the FP64 running
  // mean and variance run once per U cells instead of once per cell
  // (U <= G: every cell needs an owner lane in its group)
  constexpr int U = G < 4 ? G : 4;
  double se = 0.0, se_avg = 0.0;
  // NM modes (NM = kMaxOther + 1: nm = the run-time count, at most NM)
  const int nm = NM <= kMaxOther ? NM : a.nmodes;
  for (int q0 = blockIdx.x * cells_per_block * U; q0 < a.ntest; q0 += gridDim.x * cells_per_block * U) {
    int qs[U], ix[U][NM];
#pragma unroll
    for (int u = 0; u < U; u++) {
      qs[u] = q0 + (threadIdx.x / G) * U + u;
#pragma unroll
      for (int m = 0; m < NM; m++) ix[u][m] = (qs[u] < a.ntest && m < nm) ? a.idx[(size_t)qs[u] * nm + m] : -1;
    }
    float dot[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      dot[u] = 0.f;
      if (ix[u][0] >= 0) {
        float4 v = reinterpret_cast<const float4*>(a.f[0] + (size_t)ix[u][0] * KP)[sub];
#pragma unroll
        for (int m = 1; m < NM; m++) {
          if (m >= nm) break;
          const float4 w = reinterpret_cast<const float4*>(a.f[m] + (size_t)ix[u][m] * KP)[sub];
          v.x *= w.x;
          v.y *= w.y;
          v.z *= w.z;
          v.w *= w.w;
        }
        dot[u] = (v.x + v.y) + (v.z + v.w);
      }
    }
    float d = 0.f;
    int q = a.ntest;
#pragma unroll
    for (int u = 0; u < U; u++) {
      float du = dot[u];
#pragma unroll
      for (int off = 1; off < G; off <<= 1) du += __shfl_xor(du, off, G);
      if (sub == u) {
        d = du;
        q = qs[u];
      }
    }
    if (sub < U && q < a.ntest) {
      double p;
      if constexpr (PROBIT) {
        p = 0.5 * erfc(-(double)d * M_SQRT1_2);
        a.pred1[q] = p;
      } else {
        p = a.mean + (double)d;
      }
      double pa;
      if (n == 0) {
        pa = p;
        a.pred_var[q] = 0.0;
      } else {
        const double delta = p - a.pred[q];
        pa = a.pred[q] + delta / (n + 1);
        a.pred_var[q] += delta * (p - pa);
      }
      a.pred[q] = pa;
      const double y = a.vals[q];
      se += (y - p) * (y - p);
      se_avg += (y - pa) * (y - pa);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    se += __shfl_down(se, off, 64);
    se_avg += __shfl_down(se_avg, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = se;
    red[1][threadIdx.x >> 6] = se_avg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < kEvalThreads / 64; w++) {
      s0 += red[0][w];
      s1 += red[1][w];
    }
    a.partial[2 * blockIdx.x] = s0;
    a.partial[2 * blockIdx.x + 1] = s1;
  }
}

__global__ void __launch_bounds__(256) k_eval_finish(EvalArgs a) {
  __shared__ double red[2][256];
  double s0 = 0.0, s1 = 0.0;
  for (int b = threadIdx.x; b < a.nblk; b += 256) {
    s0 += a.partial[2 * b];
    s1 += a.partial[2 * b + 1];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      red[0][threadIdx.x] += red[0][threadIdx.x + off];
      red[1][threadIdx.x] += red[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // an empty test set gives sqrt(0 / 0) = NaN (bpmfutils.h:110-112)
    a.out[0] = a.ntest == 0 ? NAN : sqrt(red[0][0] / (double)a.ntest);
    a.out[1] = a.ntest == 0 ? NAN : sqrt(red[1][0] / (double)a.ntest);
  }
}

template <int KP>
static void launch_eval_k(const EvalArgs& a, hipStream_t s) {
  if (a.probit) {  // matrix data only (ProbitNoise::evalModel(TensorData) throws)
    hipLaunchKernelGGL((k_eval<KP, 2, true>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a);
    return;
  }
  switch (a.nmodes) {
    case 2: hipLaunchKernelGGL((k_eval<KP, 2, false>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_eval<KP, 3, false>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_eval<KP, 4, false>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_eval<KP, 5, false>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_eval<KP, kMaxOther + 1, false>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
  }
}

void launch_eval(const EvalArgs& a, hipStream_t s) {
  if (a.ntest > 0) {
    switch (a.KP) {
      case 8: launch_eval_k<8>(a, s); break;
      case 16: launch_eval_k<16>(a, s); break;
      case 32: launch_eval_k<32>(a, s); break;
      case 64: launch_eval_k<64>(a, s); break;
      default: launch_eval_k<128>(a, s); break;
    }
  }
  hipLaunchKernelGGL(k_eval_finish, dim3(1), dim3(256), 0, s, a);
}

__global__ void k_advance_iter(int* iter) {
  if (threadIdx.x == 0) *iter += 1;
}
void launch_advance_iter(int* iter, hipStream_t s) { hipLaunchKernelGGL(k_advance_iter, dim3(1), dim3(64), 0, s, iter); }

// ------------------------------------------------------------------
// RNG probes used by the parity tests
// ------------------------------------------------------------------
__global__ void k_rng_probe(uint64_t seed, uint32_t c0, uint32_t it, uint32_t tag, int n, uint32_t* bits,
                            float* z32, double* z64) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if ((j & 3) == 0) {
    uint32_t w[4];
    philox_at(seed, c0, (uint32_t)(j >> 2), it, tag, w);
    for (int q = 0; q < 4 && j + q < n; q++) bits[j + q] = w[q];
  }
  z32[j] = normal_f32(seed, c0, (uint32_t)j, it, tag);
  z64[j] = normal_f64(seed, c0, (uint32_t)j, it, tag);
}
void launch_rng_probe(uint64_t seed, uint32_t c0, uint32_t it, uint32_t tag, int n, uint32_t* bits, float* z32,
                      double* z64, hipStream_t s) {
  hipLaunchKernelGGL(k_rng_probe, dim3((n + 255) / 256), dim3(256), 0, s, seed, c0, it, tag, n, bits, z32, z64);
}

__global__ void k_gamma_probe(uint64_t seed, const double* shapes, int n, uint32_t it, uint32_t tag, double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = gamma_f64(seed, shapes[j], (uint32_t)j, it, tag);
}
void launch_gamma_probe(uint64_t seed, const double* shapes, int n, uint32_t it, uint32_t tag, double* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_gamma_probe, dim3((n + 255) / 256), dim3(256), 0, s, seed, shapes, n, it, tag, out);
}

}  // namespace mgpu
