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

// partial[0][q] = sum_b partial[b][q] in a fixed order (four interleaved
// strands of blocks combined in strand order).  Each work-group owns 64
// entries q and reads every block's copy of them before it writes row 0.
__global__ void __launch_bounds__(256) k_stats_reduce(double* partial, int nblk, int width) {
  __shared__ double red[4][64];
  const int e = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + e;
  double s = 0.0;
  if (q < width) {
    int b = g;
    for (; b + 28 < nblk; b += 32) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = partial[(size_t)(b + 4 * u) * width + q];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; b < nblk; b += 4) s += partial[(size_t)b * width + q];
  }
  red[g][e] = s;
  __syncthreads();
  if (g == 0 && q < width) partial[q] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

void launch_partial_reduce(double* partial, int nblk, int width, hipStream_t s) {
  if (nblk > 1) hipLaunchKernelGGL(k_stats_reduce, dim3((width + 63) / 64), dim3(256), 0, s, partial, nblk, width);
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
  if (nblk > 1) {
    const int width = K + K * K;
    hipLaunchKernelGGL(k_stats_reduce, dim3((width + 63) / 64), dim3(256), 0, s, partial, nblk, width);
  }
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
  if (t == 0 && !ok && a.fail) atomicMax(a.fail, 1 << 30);
}

// ------------------------------------------------------------------
// The same draw for K <= 32 on one wave, in registers.  The problem is
// padded to 32 with X_pad = diag(X, I) and C_pad = diag(C, I), and J is the
// 32-reversal; then chol(X_pad^{-1}) = diag(chol(X^{-1}), I) and
// B_pad = diag(B, I), so the leading K x K blocks are the quantities above
// and every loop has a compile-time trip count.  Lane r (r = lane % 32; the
// upper half of the wave mirrors the lower) holds row r of J X_pad J and of
// its Cholesky factor M, then column r of Z = M^{-T} J C_pad^T and of
// B = J Z, then row r of Lambda.
// ------------------------------------------------------------------
struct NW32Lds {
  alignas(16) double C[32][34];   // Bartlett factor C_pad (rows), later B (rows)
  alignas(16) double Mt[32][34];  // M^T (row i = column i of M)
  alignas(16) double col[2][32];  // Cholesky column broadcast
  alignas(16) double muc[32];     // mu_c, then mu
  alignas(16) double z[32];       // normals of the mu draw
  alignas(16) double invd[32];    // 1 / M[k][k]
};

__device__ void nw_draw32(const NWArgs& a, NW32Lds& s) {
  int lane = threadIdx.x & 63;
  opaque(lane);
  const int K = a.K;
  int r = lane & 31;
  const int it = *a.iter;
  const double N = a.N;
  const double kappa_c = a.kappa + N, nu_c = a.nu + N;
  // 1. Bartlett factor: C_ii = sqrt(chi2(nu_c - i)), C_ij ~ N(0,1) for
  //    i < j < K (WishartUnit, mvnormal.cpp:217-241), identity beyond K;
  //    the K(K+1)/2 draws are spread over the wave, column by column.
  for (int q = lane; q < 32 * 32; q += 64) {
    const int i = q >> 5, j = q & 31;
    if (i <= j && j < K) continue;
    s.C[i][j] = (i == j) ? 1.0 : 0.0;
  }
  const int ndraw = K * (K + 1) / 2;
  for (int t = lane; t < ndraw; t += 64) {
    int j = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
    while ((j + 1) * (j + 2) / 2 <= t) j++;
    while (j * (j + 1) / 2 > t) j--;
    const int i = t - j * (j + 1) / 2;
    double c;
    if (i == j)
      c = sqrt(2.0 * gamma_f64(a.seed, 0.5 * (nu_c - i), (uint32_t)i, (uint32_t)it, make_tag(TAG_WISHART_CHI, a.mode)));
    else
      c = normal_f64(a.seed, (uint32_t)i, (uint32_t)j, (uint32_t)it, make_tag(TAG_WISHART_Z, a.mode));
    s.C[i][j] = c;
  }
  if (lane < 32) {
    s.z[lane] = lane < K ? normal_f64(a.seed, 0u, (uint32_t)lane, (uint32_t)it, make_tag(TAG_NW_MU, a.mode)) : 0.0;
    // 2. mu_c (mvnormal.cpp:322-323)
    s.muc[lane] = lane < K ? (a.kappa * a.mu0[lane] + a.partial[lane]) / kappa_c : 0.0;
  }
  lds_wait();
  // 3. row r of Y = J X_pad J, X = T + S + kappa mu0 mu0^T - kappa_c mu_c mu_c^T
  //    (mvnormal.cpp:324-325; Macau adds lambda_beta beta beta^T to T,
  //    latentprior.cpp:258-260)
  const double lb = a.lambda_beta ? *a.lambda_beta : 0.0;
  const int KK = K * K;
  // the whole matrix is formed by the wave in LDS (s.Mt as staging, entry
  // (r, c) at Mt[r][c]) with branch-free loads, then each lane takes its row
  {
    const double* bbt = a.BBt_partial ? a.BBt_partial + K : a.partial + K;
    const double lbb = a.BBt_partial ? lb : 0.0;
#pragma unroll 8
    for (int q = lane; q < 32 * 32; q += 64) {
      const int rr = q >> 5, c = q & 31;
      const int xi = 31 - rr, xj = 31 - c;
      const bool in = xi < K && xj < K;
      const int xic = xi < K ? xi : 0, xjc = xj < K ? xj : 0;
      const int e = xic + xjc * K;
      const double x = a.WI[e] + lbb * bbt[e] + a.partial[K + e] + a.kappa * a.mu0[xic] * a.mu0[xjc] -
                       kappa_c * s.muc[xic] * s.muc[xjc];
      s.Mt[rr][c] = in ? x : (rr == c ? 1.0 : 0.0);
    }
  }
  (void)KK;
  lds_wait();
  double Y[32];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const double2 t2 = *reinterpret_cast<const double2*>(&s.Mt[r][2 * q]);
    Y[2 * q] = t2.x;
    Y[2 * q + 1] = t2.y;
  }
  // 4. M = chol(Y), row r on lane r (right-looking, column k + 1 published
  //    ahead of the rest of step k)
  bool bad = false;
  s.col[0][r] = Y[0];
#pragma unroll
  for (int k = 0; k < 32; k++) {
    opaque(r);
    lds_wait();
    double col[32];
#pragma unroll
    for (int q = k / 2; q < 16; q++) {
      const double2 t2 = *reinterpret_cast<const double2*>(&s.col[k & 1][2 * q]);
      col[2 * q] = t2.x;
      col[2 * q + 1] = t2.y;
    }
    const bool badp = !(col[k] > 0.0);
    bad = bad || badp;
    const double d = badp ? 1.0 : col[k];
    const double rsd = 1.0 / sqrt(d);
    s.invd[k] = rsd;  // 1 / M[k][k] (every lane stores the same value)
    const bool gt = r > k;
    const double sc = gt ? Y[k] * (rsd * rsd) : 0.0;
    if (k + 1 < 32) {
      Y[k + 1] -= sc * col[k + 1];
      s.col[(k + 1) & 1][r] = Y[k + 1];
    }
#pragma unroll
    for (int j = k + 2; j < 32; j++) Y[j] -= sc * col[j];
    Y[k] = (r >= k) ? Y[k] * rsd : 0.0;  // M[r][k]; zero above the diagonal
  }
  // M^T to LDS: row i of Mt = column i of M
#pragma unroll
  for (int c = 0; c < 32; c++) s.Mt[c][r] = Y[c];
  lds_wait();
  // 5. Z = M^{-T} R, R = J C_pad^T (R[i][c] = C_pad[c][31 - i]); lane r
  //    solves for column r of Z from the bottom row up
  double z[32];
#pragma unroll
  for (int i = 31; i >= 0; i--) {
    int o = 0;  // opaque zero offset: the row's reads stay in this step
    opaque(o);
    double acc = s.C[r][31 - i];
#pragma unroll
    for (int q = (i + 1) / 2; q < 16; q++) {
      const double2 m2 = *reinterpret_cast<const double2*>(&s.Mt[i][2 * q + o]);
      if (2 * q > i) acc -= m2.x * z[2 * q];
      if (2 * q + 1 > i) acc -= m2.y * z[2 * q + 1];
    }
    z[i] = acc * s.invd[i + o];
  }
  // 6. B = J Z: lane r holds column r of B (b[i] = z[31 - i]); rows to LDS
  lds_wait();
#pragma unroll
  for (int i = 0; i < 32; i++) s.C[i][r] = z[31 - i];
  lds_wait();
  // 7. mu = mu_c + B^{-T} z / sqrt(kappa_c) (MvNormal_prec, mvnormal.cpp:169-184):
  //    B^T x = z from the last row up; lane r holds x_r and column r of B
  {
    double acc = s.z[r];
    const double invd = 1.0 / s.C[r][r];
#pragma unroll
    for (int j = 31; j >= 0; j--) {
      opaque(r);
      const double xj = rdlane(acc * invd, j);
      acc = (r < j) ? acc - z[31 - j] * xj : (r == j ? xj : acc);
    }
    const double mu = s.muc[r] + acc / sqrt(kappa_c);
    lds_wait();
    s.muc[r] = mu;
  }
  if (lane < 32 && r < K) {
    a.mu[r] = s.muc[r];
    if (a.Lchol) {
#pragma unroll
      for (int j = 0; j < 32; j++)
        if (j < K) a.Lchol[j + r * K] = z[31 - j];  // B[j][r]
    }
  }
  lds_wait();
  // 8. Lambda = B B^T, row r, written out with Lambda mu as it is formed
  double brow[32];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const double2 t2 = *reinterpret_cast<const double2*>(&s.C[r][2 * q]);
    brow[2 * q] = t2.x;
    brow[2 * q + 1] = t2.y;
  }
  const int KP = a.KP;
  const bool wr = lane < 32;
  double lm = 0.0;
  for (int j = 0; j < 32; j++) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const double2 t2 = *reinterpret_cast<const double2*>(&s.C[j][2 * q]);
      acc += brow[2 * q] * t2.x + brow[2 * q + 1] * t2.y;
    }
    const bool in = r < K && j < K;
    if (in) lm += acc * s.muc[j];
    if (wr && in) a.Lambda[r + j * K] = acc;
    if (wr && r < KP && j < KP) a.Lambda_f[r + j * KP] = in ? (float)acc : (r == j ? 1.f : 0.f);
  }
  if (wr && r < KP) a.Lmu_f[r] = (float)lm;
  if (lane == 0 && bad && a.fail) atomicMax(a.fail, 1 << 30);
}

__global__ void __launch_bounds__(64) k_nw_draw32(NWBatch b) {
  __shared__ NW32Lds s;
  nw_draw32(b.m[blockIdx.x], s);
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
  bool small = true;
  for (int q = 0; q < b.n; q++) small = small && b.m[q].K <= 32 && b.m[q].nblk == 1;
  if (small) {
    hipLaunchKernelGGL(k_nw_draw32, dim3(b.n), dim3(64), 0, s, b);
    return;
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

template <int KP, int NM>
__global__ void __launch_bounds__(kEvalThreads) k_eval(EvalArgs a) {
  constexpr int G = KP / 4;  // lanes per test cell
  __shared__ double red[2][kEvalThreads / 64];
  const int it = *a.iter;
  const int n = it < a.burnin ? 0 : it - a.burnin;
  const int sub = threadIdx.x % G;
  const int cells_per_block = kEvalThreads / G;
  double se = 0.0, se_avg = 0.0;
  for (int q0 = blockIdx.x * cells_per_block; q0 < a.ntest; q0 += gridDim.x * cells_per_block) {
    const int q = q0 + threadIdx.x / G;
    float dot = 0.f;
    if (q < a.ntest) {
      const int* ix = a.idx + (size_t)q * NM;
      float4 v = reinterpret_cast<const float4*>(a.f[0] + (size_t)ix[0] * KP)[sub];
#pragma unroll
      for (int m = 1; m < NM; m++) {
        const float4 w = reinterpret_cast<const float4*>(a.f[m] + (size_t)ix[m] * KP)[sub];
        v.x *= w.x;
        v.y *= w.y;
        v.z *= w.z;
        v.w *= w.w;
      }
      dot = (v.x + v.y) + (v.z + v.w);
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) dot += __shfl_xor(dot, off, G);
    if (sub == 0 && q < a.ntest) {
      const double p = a.mean + (double)dot;
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
  switch (a.nmodes) {
    case 2: hipLaunchKernelGGL((k_eval<KP, 2>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_eval<KP, 3>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_eval<KP, 4>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_eval<KP, 5>), dim3(a.nblk), dim3(kEvalThreads), 0, s, a); break;
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
