/*
 * oracle.c -- FP64 CPU restatement of the jaak-s/macau Gibbs sampling step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Loop structure follows the
 * reference: OpenMP "dynamic,2" over rows for the latent samplers
 * (latentprior.cpp:73,173,222,246), scalar rank-1 Gram updates per nonzero,
 * Cholesky per row, two triangular solves.  No Eigen: the dense kernels are
 * written out here (Cholesky-Banachiewicz, forward/back substitution).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ORC_MUTANT = n > 0 builds the oracle with planted mistake n (see the
 * "planted mistake" switches below).  Only tests/test_oracle_macau.py builds
 * such variants, into a temporary directory, to show that its pins fail on
 * them; liboracle.so is always built with ORC_MUTANT = 0. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random  */
/* numbers: as easy as 1, 2, 3").  Replaces the reference's mt19937    */
/* streams (mvnormal.cpp:19-166), which are unseeded.                 */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void philox_at(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t it, uint32_t tag,
                             uint32_t w[4]) {
  uint32_t ctr[4] = {c0, c1, it, tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, w);
}

/* 24-bit uniforms: exactly representable in FP32 and FP64 */
static inline double u_open01(uint32_t x) { return ((double)(x >> 8) + 1.0) * (1.0 / 16777216.0); }
static inline double u_co01(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

/* Box-Muller (Box & Muller 1958) on the two word pairs of one block. */
static inline void box_muller4(const uint32_t w[4], double z[4]) {
  for (int p = 0; p < 2; p++) {
    double r = sqrt(-2.0 * log(u_open01(w[2 * p])));
    double th = 2.0 * M_PI * u_co01(w[2 * p + 1]);
    z[2 * p] = r * cos(th);
    z[2 * p + 1] = r * sin(th);
  }
}

void orc_normal_block(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t it, uint32_t tag, double z[4]) {
  uint32_t w[4];
  philox_at(seed, c0, c1, it, tag, w);
  box_muller4(w, z);
}

double orc_normal(uint64_t seed, uint32_t c0, uint32_t j, uint32_t it, uint32_t tag) {
  double z[4];
  orc_normal_block(seed, c0, j >> 2, it, tag, z);
  return z[j & 3];
}

void orc_normals(uint64_t seed, uint32_t c0, uint32_t n, uint32_t it, uint32_t tag, double* out) {
  double z[4];
  for (uint32_t j = 0; j < n; j++) {
    if ((j & 3) == 0) orc_normal_block(seed, c0, j >> 2, it, tag, z);
    out[j] = z[j & 3];
  }
}

/* Gamma(shape, 1) by Marsaglia & Tsang, "A simple method for generating
 * gamma variables", ACM TOMS 26(3), 2000.  Attempt t consumes block
 * philox(c0, t, it, tag): normal from (w0,w1), acceptance uniform from w2.
 * shape < 1 uses Gamma(shape+1) * U^(1/shape) with U from block
 * (c0, 0xffffffff, it, tag).  Replaces std::gamma_distribution
 * (mvnormal.cpp:163-166, 223-224). */
double orc_gamma(uint64_t seed, double shape, uint32_t c0, uint32_t it, uint32_t tag) {
  double a = shape < 1.0 ? shape + 1.0 : shape;
  double d = a - 1.0 / 3.0;
  double c = 1.0 / sqrt(9.0 * d);
  double g = d;
  for (uint32_t t = 0; t < 32; t++) {
    uint32_t w[4];
    philox_at(seed, c0, t, it, tag, w);
    double x = sqrt(-2.0 * log(u_open01(w[0]))) * cos(2.0 * M_PI * u_co01(w[1]));
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    double u = u_open01(w[2]);
    double x2 = x * x;
    if (u < 1.0 - 0.0331 * x2 * x2 || log(u) < 0.5 * x2 + d * (1.0 - v + log(v))) {
      g = d * v;
      break;
    }
  }
  if (shape < 1.0) {
    uint32_t w[4];
    philox_at(seed, c0, 0xffffffffu, it, tag, w);
    g *= pow(u_open01(w[0]), 1.0 / shape);
  }
  return g;
}

/* ------------------------------------------------------------------ */
/* Dense helpers (replacing Eigen::LLT and triangularView solves)     */
/* ------------------------------------------------------------------ */
int orc_chol_lower(int n, double* A) {
  for (int j = 0; j < n; j++) {
    double d = A[j + (size_t)j * n];
    for (int k = 0; k < j; k++) d -= A[j + (size_t)k * n] * A[j + (size_t)k * n];
    if (!(d > 0.0)) return -1;
    d = sqrt(d);
    A[j + (size_t)j * n] = d;
    for (int i = j + 1; i < n; i++) {
      double s = A[i + (size_t)j * n];
      for (int k = 0; k < j; k++) s -= A[i + (size_t)k * n] * A[j + (size_t)k * n];
      A[i + (size_t)j * n] = s / d;
    }
    for (int i = 0; i < j; i++) A[i + (size_t)j * n] = 0.0;
  }
  return 0;
}

void orc_trsv_lower(int n, const double* L, double* x) {
  for (int i = 0; i < n; i++) {
    double s = x[i];
    for (int k = 0; k < i; k++) s -= L[i + (size_t)k * n] * x[k];
    x[i] = s / L[i + (size_t)i * n];
  }
}

void orc_trsv_upper_t(int n, const double* L, double* x) {
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int k = i + 1; k < n; k++) s -= L[k + (size_t)i * n] * x[k];
    x[i] = s / L[i + (size_t)i * n];
  }
}

/* inverse of an SPD matrix through its Cholesky factor (replaces X.inverse(),
 * mvnormal.cpp:326; equal up to rounding for the SPD X built there) */
static int spd_inverse(int n, const double* X, double* Xinv) {
  double* L = (double*)malloc(sizeof(double) * n * n);
  memcpy(L, X, sizeof(double) * n * n);
  if (orc_chol_lower(n, L)) { free(L); return -1; }
  double* e = (double*)malloc(sizeof(double) * n);
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < n; i++) e[i] = (i == j) ? 1.0 : 0.0;
    orc_trsv_lower(n, L, e);
    orc_trsv_upper_t(n, L, e);
    for (int i = 0; i < n; i++) Xinv[i + (size_t)j * n] = e[i];
  }
  free(e);
  free(L);
  return 0;
}

/* ------------------------------------------------------------------ */
/* a6/a7/a8 per-row latent samplers                                   */
/* ------------------------------------------------------------------ */
int64_t orc_sample_latents(int K, const orc_rows* rows, double mean, const double* const* others,
                           double alpha, const double* mu, const double* Lambda,
                           const double* mu_rows, uint64_t seed, uint32_t it, int mode, double* U) {
  const int64_t N = rows->nrows;
  const int no = rows->nother;
  int64_t failed = 0;
  /* Lambda * mu is shared across rows for BPMFPrior (latentprior.cpp:378) */
#pragma omp parallel
  {
    double* MM = (double*)malloc(sizeof(double) * K * K);
    double* rr = (double*)malloc(sizeof(double) * K);
    double* col = (double*)malloc(sizeof(double) * K);
#pragma omp for schedule(dynamic, 2)
    for (int64_t n = 0; n < N; n++) {
      memcpy(MM, Lambda, sizeof(double) * K * K);
      for (int k = 0; k < K; k++) rr[k] = 0.0;
      for (int64_t e = rows->row_ptr[n]; e < rows->row_ptr[n + 1]; e++) {
        const double* c0 = others[0] + (size_t)rows->idx[e * no] * K;
        for (int k = 0; k < K; k++) col[k] = c0[k];
        for (int t = 1; t < no; t++) {
          const double* ct = others[t] + (size_t)rows->idx[e * no + t] * K;
          for (int k = 0; k < K; k++) col[k] *= ct[k];
        }
        const double ra = (rows->vals[e] - mean) * alpha;
        /* MM.triangularView<Lower>() += alpha * col * col^T (latentprior.cpp:369) */
        for (int j = 0; j < K; j++) {
          const double aj = alpha * col[j];
          for (int i = j; i < K; i++) MM[i + (size_t)j * K] += col[i] * aj;
        }
        for (int k = 0; k < K; k++) rr[k] += col[k] * ra;
      }
      if (orc_chol_lower(K, MM)) {
#pragma omp atomic write
        failed = n + 1;
        continue;
      }
      const double* m = mu_rows ? mu_rows + (size_t)n * K : mu;
      for (int i = 0; i < K; i++) {
        double s = 0.0;
        for (int j = 0; j < K; j++) s += Lambda[i + (size_t)j * K] * m[j];
        rr[i] += s;
      }
      orc_trsv_lower(K, MM, rr);
      double z[4];
      for (int k = 0; k < K; k++) {
        if ((k & 3) == 0) orc_normal_block(seed, (uint32_t)n, (uint32_t)(k >> 2), it,
                                           ORC_TAG(ORC_TAG_LATENT, mode), z);
        rr[k] += z[k & 3];
      }
      orc_trsv_upper_t(K, MM, rr);
      memcpy(U + (size_t)n * K, rr, sizeof(double) * K);
    }
    free(MM);
    free(rr);
    free(col);
  }
  return failed ? -failed : 0;
}

/* ------------------------------------------------------------------ */
/* a9 Normal-Wishart hyperprior                                        */
/* ------------------------------------------------------------------ */
int orc_normal_wishart_from_stats(int K, int64_t N, const double* NU, const double* NS,
                                  const double* mu0, double kappa, const double* T, double nu,
                                  uint64_t seed, uint32_t it, int mode, double* mu_out,
                                  double* Lambda_out) {
  const size_t KK = (size_t)K * K;
  double* X = (double*)malloc(sizeof(double) * KK);
  double* Tc = (double*)malloc(sizeof(double) * KK);
  double* L = (double*)malloc(sizeof(double) * KK);
  double* C = (double*)calloc(KK, sizeof(double));
  double* AU = (double*)malloc(sizeof(double) * KK);
  double* tmp = (double*)malloc(sizeof(double) * KK);
  double* mu_c = (double*)malloc(sizeof(double) * K);
  double* z = (double*)malloc(sizeof(double) * K);
  int rc = 0;
  /* mvnormal.cpp:322-326 */
  const double nu_c = nu + (double)N;
  const double kappa_c = kappa + (double)N;
  for (int k = 0; k < K; k++) mu_c[k] = (kappa * mu0[k] + NU[k]) / (kappa + (double)N);
  for (int j = 0; j < K; j++)
    for (int i = 0; i < K; i++)
      X[i + (size_t)j * K] = T[i + (size_t)j * K] + NS[i + (size_t)j * K] +
                             kappa * mu0[i] * mu0[j] - kappa_c * mu_c[i] * mu_c[j];
  if (spd_inverse(K, X, Tc)) { rc = -1; goto done; }
  /* Wishart(T_c, nu_c) (mvnormal.cpp:243-266): L = chol(T_c) */
  memcpy(L, Tc, sizeof(double) * KK);
  if (orc_chol_lower(K, L)) { rc = -2; goto done; }
  /* WishartUnit (mvnormal.cpp:217-241): upper C, C_ii = sqrt(chi2(df - i)) */
  for (int i = 0; i < K; i++) {
    double g = orc_gamma(seed, 0.5 * (nu_c - i), (uint32_t)i, it, ORC_TAG(ORC_TAG_WISHART_CHI, mode));
    C[i + (size_t)i * K] = sqrt(2.0 * g);
    for (int j = i + 1; j < K; j++)
      C[i + (size_t)j * K] = orc_normal(seed, (uint32_t)i, (uint32_t)j, it, ORC_TAG(ORC_TAG_WISHART_Z, mode));
  }
  /* AU = C^T C */
  for (int j = 0; j < K; j++)
    for (int i = 0; i < K; i++) {
      double s = 0.0;
      for (int k = 0; k < K; k++) s += C[k + (size_t)i * K] * C[k + (size_t)j * K];
      AU[i + (size_t)j * K] = s;
    }
  /* Lambda = L AU L^T */
  for (int j = 0; j < K; j++)
    for (int i = 0; i < K; i++) {
      double s = 0.0;
      for (int k = 0; k < K; k++) s += L[i + (size_t)k * K] * AU[k + (size_t)j * K];
      tmp[i + (size_t)j * K] = s;
    }
  for (int j = 0; j < K; j++)
    for (int i = 0; i < K; i++) {
      double s = 0.0;
      for (int k = 0; k < K; k++) s += tmp[i + (size_t)k * K] * L[j + (size_t)k * K];
      Lambda_out[i + (size_t)j * K] = s;
    }
  /* NormalWishart: mu = MvNormal_prec(Lambda * kappa_c, mu_c)
   * (mvnormal.cpp:169-184,273): mu_c + chol(kappa_c Lambda)^{-T} z */
  for (size_t q = 0; q < KK; q++) tmp[q] = Lambda_out[q] * kappa_c;
  if (orc_chol_lower(K, tmp)) { rc = -3; goto done; }
  orc_normals(seed, 0, (uint32_t)K, it, ORC_TAG(ORC_TAG_NW_MU, mode), z);
  orc_trsv_upper_t(K, tmp, z);
  for (int k = 0; k < K; k++) mu_out[k] = mu_c[k] + z[k];
done:
  free(X); free(Tc); free(L); free(C); free(AU); free(tmp); free(mu_c); free(z);
  return rc;
}

int orc_cond_normal_wishart(int K, int64_t N, const double* U, const double* mu0, double kappa,
                            const double* T, double nu, uint64_t seed, uint32_t it, int mode,
                            double* mu_out, double* Lambda_out) {
  /* NU = U.rowwise().sum(); NS = U U^T (mvnormal.cpp:319-320) */
  double* NU = (double*)calloc(K, sizeof(double));
  double* NS = (double*)calloc((size_t)K * K, sizeof(double));
  for (int64_t n = 0; n < N; n++) {
    const double* u = U + (size_t)n * K;
    for (int j = 0; j < K; j++) {
      NU[j] += u[j];
      for (int i = 0; i < K; i++) NS[i + (size_t)j * K] += u[i] * u[j];
    }
  }
  int rc = orc_normal_wishart_from_stats(K, N, NU, NS, mu0, kappa, T, nu, seed, it, mode, mu_out, Lambda_out);
  free(NU);
  free(NS);
  return rc;
}

/* ------------------------------------------------------------------ */
/* a12 noise, a13 eval                                                 */
/* ------------------------------------------------------------------ */
double orc_train_sse(int K, int nmodes, const orc_rows* rows0, double mean, const double* const* U) {
  /* AdaptiveGaussianNoise::update (noisemodels.cpp:55-81, 83-124) */
  /* per-row sums combined in row order: independent of the thread count */
  const int no = rows0->nother;
  double* per_row = (double*)calloc((size_t)rows0->nrows + 1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t n = 0; n < rows0->nrows; n++) {
    const double* u = U[0] + (size_t)n * K;
    double s = 0.0;
    for (int64_t e = rows0->row_ptr[n]; e < rows0->row_ptr[n + 1]; e++) {
      double yhat = mean;
      for (int d = 0; d < K; d++) {
        double t = u[d];
        for (int m = 1; m < nmodes; m++) t *= U[m][(size_t)rows0->idx[e * no + m - 1] * K + d];
        yhat += t;
      }
      double r = yhat - rows0->vals[e];
      s += r * r;
    }
    per_row[n] = s;
  }
  double sumsq = 0.0;
  for (int64_t n = 0; n < rows0->nrows; n++) sumsq += per_row[n];
  free(per_row);
  return sumsq;
}

double orc_noise_alpha(double sse, int64_t nnz, double var_total, int is_tensor, double alpha_max,
                       uint64_t seed, uint32_t it) {
  const double a0 = 0.5;
  const double b0 = is_tensor ? 0.5 : 0.5 * var_total; /* noisemodels.cpp:73-76 vs 117-118 */
  const double aN = a0 + (double)nnz / 2.0;
  const double bN = b0 + sse / 2.0;
  double alpha = orc_gamma(seed, aN, 0, it, ORC_TAG(ORC_TAG_NOISE, 0)) / bN; /* rgamma(aN, 1/bN) */
  if (alpha > alpha_max) alpha = alpha_max;
  return alpha;
}

void orc_eval_rmse(int K, int nmodes, int64_t ntest, const int32_t* test_idx, const double* test_vals,
                   double mean, const double* const* U, int n, double* pred, double* pred_var,
                   double* rmse1, double* rmse_avg) {
  /* per-cell errors summed in cell order: independent of the thread count */
  double* e1 = (double*)malloc(sizeof(double) * ((size_t)ntest + 1));
  double* ea = (double*)malloc(sizeof(double) * ((size_t)ntest + 1));
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < ntest; t++) {
    const int32_t* ix = test_idx + t * nmodes;
    double p = mean;
    for (int d = 0; d < K; d++) {
      double v = U[0][(size_t)ix[0] * K + d];
      for (int m = 1; m < nmodes; m++) v *= U[m][(size_t)ix[m] * K + d];
      p += v;
    }
    double pa;
    if (n == 0) {
      pa = p;
    } else {
      double delta = p - pred[t];
      pa = pred[t] + delta / (n + 1);
      pred_var[t] += delta * (p - pa);
    }
    e1[t] = (test_vals[t] - p) * (test_vals[t] - p);
    ea[t] = (test_vals[t] - pa) * (test_vals[t] - pa);
    pred[t] = pa;
  }
  double se = 0.0, se_avg = 0.0;
  for (int64_t t = 0; t < ntest; t++) {
    se += e1[t];
    se_avg += ea[t];
  }
  free(e1);
  free(ea);
  /* matrix: sqrt(se/N) is NaN for an empty test set (bpmfutils.h:110-112);
   * tensor: explicit NaN (bpmfutils.cpp:39-43) -- identical result */
  *rmse1 = sqrt(se / (double)ntest);
  *rmse_avg = sqrt(se_avg / (double)ntest);
}

/* ------------------------------------------------------------------ */
/* a10/a11 Macau link matrix: SpMM, block CG, lambda_beta              */
/* ------------------------------------------------------------------ */
void orc_csr_A_mul_Bn(int nrhs, const orc_csr* A, const double* B, double* out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t r = 0; r < A->nrow; r++) {
    double* o = out + (size_t)r * nrhs;
    for (int j = 0; j < nrhs; j++) o[j] = 0.0;
    for (int64_t e = A->row_ptr[r]; e < A->row_ptr[r + 1]; e++) {
      const double* b = B + (size_t)A->cols[e] * nrhs;
      const double v = A->vals ? A->vals[e] : 1.0;
      for (int j = 0; j < nrhs; j++) o[j] += b[j] * v;
    }
  }
}

void orc_AtA_mul_B(int nrhs, const orc_csr* A, const orc_csr* At, double reg, const double* B, double* out) {
  double* inner = (double*)malloc(sizeof(double) * (size_t)nrhs * A->nrow);
  orc_csr_A_mul_Bn(nrhs, A, B, inner);
  orc_csr_A_mul_Bn(nrhs, At, inner, out);
  const size_t n = (size_t)nrhs * A->ncol;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; q++) out[q] += reg * B[q];
  free(inner);
}

/* lower triangle of A B^T for nrhs x nfeat operands stored K x nfeat
 * (A_mul_Bt_omp_sym / A_mul_At_omp, linop.cpp:271-372), then symmetrised */
static void rows_dot(int nr, int64_t nfeat, const double* A, const double* B, double* out) {
  /* fixed 4096-feature chunks combined in chunk order: independent of the
   * thread count */
  const int64_t CH = 4096;
  const int64_t nch = (nfeat + CH - 1) / CH;
  const size_t NN = (size_t)nr * nr;
  double* part = (double*)calloc((size_t)(nch > 0 ? nch : 1) * NN, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t c = 0; c < nch; c++) {
    double* loc = part + (size_t)c * NN;
    const int64_t f1 = (c + 1) * CH < nfeat ? (c + 1) * CH : nfeat;
    for (int64_t f = c * CH; f < f1; f++) {
      const double* a = A + (size_t)f * nr;
      const double* b = B + (size_t)f * nr;
      for (int j = 0; j < nr; j++)
        for (int i = j; i < nr; i++) loc[i + j * nr] += a[i] * b[j];
    }
  }
  for (size_t q = 0; q < NN; q++) out[q] = 0.0;
  for (int64_t c = 0; c < nch; c++)
    for (size_t q = 0; q < NN; q++) out[q] += part[(size_t)c * NN + q];
  free(part);
  for (int j = 0; j < nr; j++)
    for (int i = 0; i < j; i++) out[i + j * nr] = out[j + i * nr];
}

/* Eigen's unblocked LLT (left-looking): factors column k only while the
 * pivot x = a_kk - |L(k,0:k)|^2 is positive; on the first non-positive pivot
 * it stops, leaving columns >= k as the input's lower triangle, and LLT::solve
 * then uses that partially factored matrix.  The reference's block CG relies
 * on this when the block Krylov space is exhausted (tests.cpp:183-207: three
 * right-hand sides on a 4-dimensional system converge in two iterations). */
static int eigen_llt(int n, double* A) {
  for (int k = 0; k < n; k++) {
    double x = A[k + k * n];
    for (int j = 0; j < k; j++) x -= A[k + j * n] * A[k + j * n];
    if (x <= 0.0) return k;
    x = sqrt(x);
    A[k + k * n] = x;
    for (int i = k + 1; i < n; i++) {
      double s = A[i + k * n];
      for (int j = 0; j < k; j++) s -= A[i + j * n] * A[k + j * n];
      A[i + k * n] = s / x;
    }
  }
  return -1;
}

/* X = M^{-1} Rhs with M's LLT as Eigen computes it (see eigen_llt) */
static void spd_solve(int nr, const double* M, const double* Rhs, double* X) {
  double* L = (double*)malloc(sizeof(double) * nr * nr);
  memcpy(L, M, sizeof(double) * nr * nr);
  eigen_llt(nr, L);
  for (int j = 0; j < nr; j++)
    for (int i = 0; i < j; i++) L[i + j * nr] = 0.0;
  for (int j = 0; j < nr; j++) {
    double* x = X + (size_t)j * nr;
    for (int i = 0; i < nr; i++) x[i] = Rhs[i + (size_t)j * nr];
    orc_trsv_lower(nr, L, x);
    orc_trsv_upper_t(nr, L, x);
  }
  free(L);
}

int orc_blockcg(int nrhs, const orc_csr* A, const orc_csr* At, double reg, const double* B, double tol, double* X) {
  const int64_t nfeat = A->ncol;
  const size_t sz = (size_t)nrhs * nfeat;
  const double tolsq = tol * tol;
  double* norms = (double*)malloc(sizeof(double) * nrhs);
  double* inorms = (double*)malloc(sizeof(double) * nrhs);
  double* R = (double*)malloc(sizeof(double) * sz);
  double* P = (double*)malloc(sizeof(double) * sz);
  double* KP = (double*)malloc(sizeof(double) * sz);
  double* RtR = (double*)malloc(sizeof(double) * nrhs * nrhs);
  double* RtR2 = (double*)malloc(sizeof(double) * nrhs * nrhs);
  double* PtKP = (double*)malloc(sizeof(double) * nrhs * nrhs);
  double* Am = (double*)malloc(sizeof(double) * nrhs * nrhs);
  double* Psi = (double*)malloc(sizeof(double) * nrhs * nrhs);
  /* linop.h:240-260: normalise each RHS row */
  for (int r = 0; r < nrhs; r++) {
    double s = 0.0;
    for (int64_t f = 0; f < nfeat; f++) s += B[(size_t)f * nrhs + r] * B[(size_t)f * nrhs + r];
    norms[r] = sqrt(s);
    inorms[r] = 1.0 / norms[r];
  }
  for (int64_t f = 0; f < nfeat; f++)
    for (int r = 0; r < nrhs; r++) {
      size_t q = (size_t)f * nrhs + r;
      X[q] = 0.0;
      R[q] = B[q] * inorms[r];
      P[q] = R[q];
    }
  rows_dot(nrhs, nfeat, R, R, RtR);
  int iter;
  for (iter = 0; iter < 100000; iter++) {
    orc_AtA_mul_B(nrhs, A, At, reg, P, KP);              /* linop.h:282 */
    rows_dot(nrhs, nfeat, P, KP, PtKP);                  /* linop.h:286 */
    /* A = (PtKP^{-1} RtR)^T (linop.h:287-289) */
    double* S = (double*)malloc(sizeof(double) * nrhs * nrhs);
    spd_solve(nrhs, PtKP, RtR, S);
    for (int j = 0; j < nrhs; j++)
      for (int i = 0; i < nrhs; i++) Am[i + j * nrhs] = S[j + i * nrhs];
    free(S);
    /* X += A P; R -= A KP (linop.h:292-300) */
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nfeat; f++) {
      double* x = X + (size_t)f * nrhs;
      double* r = R + (size_t)f * nrhs;
      const double* p = P + (size_t)f * nrhs;
      const double* kp = KP + (size_t)f * nrhs;
      for (int i = 0; i < nrhs; i++) {
        double sx = 0.0, sr = 0.0;
        for (int k = 0; k < nrhs; k++) {
          sx += Am[i + k * nrhs] * p[k];
          sr += Am[i + k * nrhs] * kp[k];
        }
        x[i] += sx;
        r[i] -= sr;
      }
    }
    rows_dot(nrhs, nfeat, R, R, RtR2);                   /* linop.h:304-305 */
    int conv = 1;
    for (int i = 0; i < nrhs; i++)
      if (!(RtR2[i + i * nrhs] < tolsq)) conv = 0;
    if (conv) break;                                     /* linop.h:309 */
    /* Psi = (RtR^{-1} RtR2)^T; P = R + Psi P (linop.h:313-326) */
    S = (double*)malloc(sizeof(double) * nrhs * nrhs);
    spd_solve(nrhs, RtR, RtR2, S);
    for (int j = 0; j < nrhs; j++)
      for (int i = 0; i < nrhs; i++) Psi[i + j * nrhs] = S[j + i * nrhs];
    free(S);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nfeat; f++) {
      double* p = P + (size_t)f * nrhs;
      const double* r = R + (size_t)f * nrhs;
      double t[64];
      for (int i = 0; i < nrhs; i++) {
        double s = 0.0;
        for (int k = 0; k < nrhs; k++) s += Psi[i + k * nrhs] * p[k];
        t[i] = s;
      }
      for (int i = 0; i < nrhs; i++) p[i] = r[i] + t[i];
    }
    double* sw = RtR; RtR = RtR2; RtR2 = sw;            /* linop.h:329 */
  }
  for (int64_t f = 0; f < nfeat; f++)
    for (int r = 0; r < nrhs; r++) X[(size_t)f * nrhs + r] *= norms[r];
  free(norms); free(inorms); free(R); free(P); free(KP);
  free(RtR); free(RtR2); free(PtKP); free(Am); free(Psi);
  return iter;
}

int orc_blockcg_blocked(int nrhs, const orc_csr* A, const orc_csr* At, double reg, const double* B,
                        double tol, int blocksize, int excess, double* X) {
  const int64_t nfeat = A->ncol;
  if (nrhs <= excess + blocksize) return orc_blockcg(nrhs, A, At, reg, B, tol, X);
  int last = 0;
  /* linop.h:214-225, including its re-solve of the tail when the loop
   * steps past a block that already absorbed the excess rows */
  for (int i = 0; i < nrhs; i += blocksize) {
    int nr = blocksize;
    if (i + blocksize + excess >= nrhs) nr = nrhs - i;
    double* Bb = (double*)malloc(sizeof(double) * nr * nfeat);
    double* Xb = (double*)malloc(sizeof(double) * nr * nfeat);
    for (int64_t f = 0; f < nfeat; f++)
      for (int r = 0; r < nr; r++) Bb[(size_t)f * nr + r] = B[(size_t)f * nrhs + i + r];
    last = orc_blockcg(nr, A, At, reg, Bb, tol, Xb);
    for (int64_t f = 0; f < nfeat; f++)
      for (int r = 0; r < nr; r++) X[(size_t)f * nrhs + i + r] = Xb[(size_t)f * nr + r];
    free(Bb);
    free(Xb);
  }
  return last;
}

void orc_posterior_lambda_beta(int K, int64_t nfeat, const double* beta, const double* Lambda, double nu,
                               double mu, double* shape, double* scale) {
  /* latentprior.cpp:317-326 */
  double* BB = (double*)malloc(sizeof(double) * K * K);
  rows_dot(K, nfeat, beta, beta, BB);
  double tr = 0.0;
  for (int i = 0; i < K; i++)
    for (int j = 0; j < K; j++) tr += BB[i + j * K] * Lambda[j + i * K];
  const double nux = nu + (double)K * (double)nfeat;
  const double mux = mu * nux / (nu + mu * tr);
  *shape = nux / 2.0;
  *scale = 2.0 * mux / nux;
  free(BB);
}

/* MvNormal_prec_omp(Lambda, n) (mvnormal.cpp:186-197) with draws of stream
 * tag: column c gets normals (c0 = c, j = 0..K-1); out = chol(Lambda)^{-T} z */
static void mvnormal_prec_cols(int K, const double* Lchol, int64_t n, uint64_t seed, uint32_t it,
                               uint32_t tag, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; c++) {
    double* o = out + (size_t)c * K;
    orc_normals(seed, (uint32_t)c, (uint32_t)K, it, tag, o);
#if ORC_MUTANT == 3 /* planted mistake (tests only): covariance Lambda instead of Lambda^{-1} */
    for (int i = K - 1; i >= 0; i--) {
      double t = 0.0;
      for (int k = 0; k <= i; k++) t += Lchol[i + (size_t)k * K] * o[k];
      o[i] = t;
    }
#else
    orc_trsv_upper_t(K, Lchol, o);
#endif
  }
}

/* At_mul_A(SparseDoubleFeat) (linop.cpp:93-118): for every feature f1, the
 * rows of F holding it, and for each such row every feature f2 >= f1 of it */
void orc_at_mul_a(const orc_csr* A, const orc_csr* At, double* out) {
  const int64_t n = A->ncol;
  memset(out, 0, sizeof(double) * (size_t)n * (size_t)n);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t f1 = 0; f1 < n; f1++) {
    for (int64_t i = At->row_ptr[f1]; i < At->row_ptr[f1 + 1]; i++) {
      const int64_t r = At->cols[i];
      const double v1 = At->vals ? At->vals[i] : 1.0;
      for (int64_t j = A->row_ptr[r]; j < A->row_ptr[r + 1]; j++) {
        const int64_t f2 = A->cols[j];
        if (f1 <= f2) out[f2 + (size_t)f1 * n] += (A->vals ? A->vals[j] : 1.0) * v1;
      }
    }
  }
}

/* K = FtF (lower) + reg I; chol_decomp(K); chol_solve_t(K, B): every row of
 * B (a latent component over the features) is solved against K */
int orc_ftf_solve(int nrhs, int64_t n, const double* FtF, double reg, const double* B, double* X) {
  double* L = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
  for (int64_t j = 0; j < n; j++)
    for (int64_t i = 0; i < n; i++) L[i + (size_t)j * n] = i >= j ? FtF[i + (size_t)j * n] : 0.0;
  for (int64_t i = 0; i < n; i++) L[i + (size_t)i * n] += reg;
  if (orc_chol_lower((int)n, L)) {
    free(L);
    return -1;
  }
  double* x = (double*)malloc(sizeof(double) * (size_t)n);
  for (int k = 0; k < nrhs; k++) {
    for (int64_t f = 0; f < n; f++) x[f] = B[k + (size_t)f * nrhs];
    orc_trsv_lower((int)n, L, x);
    orc_trsv_upper_t((int)n, L, x);
    for (int64_t f = 0; f < n; f++) X[k + (size_t)f * nrhs] = x[f];
  }
  free(x);
  free(L);
  return 0;
}

/* MacauPrior::sample_beta (latentprior.cpp:272-293): draws the link matrix
 * beta (K x nfeat) at fixed U, mu, Lambda and lambda_beta.  FtF (non-null)
 * selects the F^T F branch, else the block CG.  Returns 0, -10 when Lambda is
 * not positive definite, -11 when F^T F + lambda_beta I is not. */
int orc_sample_beta(int K, int mode, const double* U, int64_t N, const double* mu, const double* Lambda,
                    const orc_csr* F, const orc_csr* Ft, const double* FtF, double lambda_beta, double tol,
                    uint64_t seed, uint32_t it, double* beta, int* cg_iters) {
  const int64_t nfeat = F->ncol;
  const size_t KK = (size_t)K * K;
  int rc = 0;
  double* Lc = (double*)malloc(sizeof(double) * KK);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)K * N);
  double* Fty = (double*)malloc(sizeof(double) * (size_t)K * nfeat);
  double* noise = (double*)malloc(sizeof(double) * (size_t)K * nfeat);
  memcpy(Lc, Lambda, sizeof(double) * KK);
  if (orc_chol_lower(K, Lc)) { rc = -10; goto out; }
  /* tmp = U + MvNormal_prec_omp(Lambda, N) - mu (latentprior.cpp:277) */
  mvnormal_prec_cols(K, Lc, N, seed, it, ORC_TAG(ORC_TAG_BETA_U, mode), tmp);
  for (int64_t n = 0; n < N; n++)
    for (int k = 0; k < K; k++) {
#if ORC_MUTANT == 2 /* planted mistake (tests only): the mean is not subtracted */
      tmp[(size_t)n * K + k] += U[(size_t)n * K + k];
#else
      tmp[(size_t)n * K + k] += U[(size_t)n * K + k] - mu[k];
#endif
    }
  /* Ft_y = tmp F + sqrt(lambda_beta) MvNormal_prec_omp(Lambda, nfeat) (:278) */
  orc_csr_A_mul_Bn(K, Ft, tmp, Fty);
  mvnormal_prec_cols(K, Lc, nfeat, seed, it, ORC_TAG(ORC_TAG_BETA_F, mode), noise);
  {
#if ORC_MUTANT == 1 /* planted mistake (tests only): sqrt(lambda_beta) dropped */
    const double sl = 1.0;
#else
    const double sl = sqrt(lambda_beta);
#endif
    for (size_t q = 0; q < (size_t)K * nfeat; q++) Fty[q] += sl * noise[q];
  }
  if (FtF) {
    /* K = FtF + lambda_beta I, chol_decomp, chol_solve_t (:280-288) */
    if (orc_ftf_solve(K, nfeat, FtF, lambda_beta, Fty, beta)) { rc = -11; goto out; }
    *cg_iters = 0;
  } else {
    /* solve_blockcg(beta, F, lambda_beta, Ft_y, tol, 32, 8) (:291) */
    *cg_iters = orc_blockcg_blocked(K, F, Ft, lambda_beta, Fty, tol, 32, 8, beta);
  }
out:
  free(Lc); free(tmp); free(Fty); free(noise);
  return rc;
}

/* MacauPrior::update_prior (latentprior.cpp:254-264): the residual U - Uhat,
 * the Normal-Wishart draw with WI + lambda_beta beta beta^T and df + nfeat,
 * sample_beta with the new (mu, Lambda), Uhat = beta F^T and lambda_beta from
 * its gamma posterior.  Updates mu, Lambda, beta, Uhat and *lambda_beta. */
int orc_macau_update_prior(int K, int mode, const double* U, int64_t N, const orc_csr* F, const orc_csr* Ft,
                           const double* FtF, const double* WI, int df, double* beta, double* Uhat,
                           double* lambda_beta, double tol, double* mu, double* Lambda, uint64_t seed, uint32_t it,
                           int* cg_iters) {
  const int64_t nfeat = F->ncol;
  const size_t KK = (size_t)K * K;
  int rc = 0;
#if ORC_MUTANT == 6 /* planted mistake (tests only): the prior sees U, not the residual U - Uhat */
  for (size_t q = 0; q < (size_t)K * N; q++) Uhat[q] = U[q];
#else
  for (size_t q = 0; q < (size_t)K * N; q++) Uhat[q] = U[q] - Uhat[q];
#endif
  double* BBt = (double*)malloc(sizeof(double) * KK);
  double* T = (double*)malloc(sizeof(double) * KK);
  double* mu0 = (double*)calloc(K, sizeof(double));
  rows_dot(K, nfeat, beta, beta, BBt);
#if ORC_MUTANT == 4 /* planted mistake (tests only): WI + beta beta^T without lambda_beta */
  for (size_t q = 0; q < KK; q++) T[q] = WI[q] + BBt[q];
#else
  for (size_t q = 0; q < KK; q++) T[q] = WI[q] + (*lambda_beta) * BBt[q];
#endif
  rc = orc_cond_normal_wishart(K, N, Uhat, mu0, 2.0, T, (double)df + (double)nfeat, seed, it, mode, mu, Lambda);
  if (rc) goto out;
  rc = orc_sample_beta(K, mode, U, N, mu, Lambda, F, Ft, FtF, *lambda_beta, tol, seed, it, beta, cg_iters);
  if (rc) goto out;
  /* compute_uhat: Uhat = beta F^T (linop.h:196-198) */
  orc_csr_A_mul_Bn(K, F, beta, Uhat);
  {
    double shape, scale;
    orc_posterior_lambda_beta(K, nfeat, beta, Lambda, 1e-3, 1.0, &shape, &scale);
    *lambda_beta = orc_gamma(seed, shape, 0, it, ORC_TAG(ORC_TAG_LAMBDA_BETA, mode)) * scale;
  }
out:
  free(BBt); free(T); free(mu0);
  return rc;
}

/* ------------------------------------------------------------------ */
/* univariate prior MacauOnePrior                                           */
/* ------------------------------------------------------------------ */
void orc_sample_latents_uni(int K, const orc_rows* rows, double mean, const double* const* others, double alpha,
                            const double* lambda, const double* mu_rows, uint64_t seed, uint32_t it, int mode,
                            double* U) {
  const int64_t N = rows->nrows;
  const int no = rows->nother;
#pragma omp parallel
  {
    double* Q = (double*)malloc(sizeof(double) * K);
    double* z = (double*)malloc(sizeof(double) * (K + 4));
    int64_t cap = 0;
    double *yhat = NULL, *prod = NULL;
#pragma omp for schedule(dynamic, 4)
    for (int64_t n = 0; n < N; n++) {
      const int64_t beg = rows->row_ptr[n], nnz = rows->row_ptr[n + 1] - beg;
      if (nnz > cap) {
        cap = nnz;
        yhat = (double*)realloc(yhat, sizeof(double) * (size_t)cap);
        prod = (double*)realloc(prod, sizeof(double) * (size_t)cap * K);
      }
      double* u = U + (size_t)n * K;
      /* Qi = lambda + alpha sum v^2, Yhat = mean + u . v (:60-66, :137-148) */
      for (int d = 0; d < K; d++) Q[d] = lambda[d];
      for (int64_t j = 0; j < nnz; j++) {
        const int64_t e = beg + j;
        double* v = prod + (size_t)j * K;
        const double* c0 = others[0] + (size_t)rows->idx[e * no] * K;
        for (int d = 0; d < K; d++) v[d] = c0[d];
        for (int t = 1; t < no; t++) {
          const double* ct = others[t] + (size_t)rows->idx[e * no + t] * K;
          for (int d = 0; d < K; d++) v[d] *= ct[d];
        }
        double dot = 0.0;
        for (int d = 0; d < K; d++) {
          Q[d] += alpha * v[d] * v[d];
          dot += u[d] * v[d];
        }
        yhat[j] = mean + dot;
      }
      for (int d = 0; d < K; d += 4) orc_normal_block(seed, (uint32_t)n, (uint32_t)(d >> 2), it,
                                                      ORC_TAG(ORC_TAG_LATENT, mode), z + d);
      /* coordinate updates (:70-93, :154-185) */
      for (int d = 0; d < K; d++) {
        const double uid = u[d];
        double L = lambda[d] * mu_rows[(size_t)n * K + d];
        for (int64_t j = 0; j < nnz; j++) {
          const double vjd = prod[(size_t)j * K + d];
          L += alpha * (rows->vals[beg + j] - (yhat[j] - uid * vjd)) * vjd;
        }
        const double var = 1.0 / Q[d];
        u[d] = L * var + sqrt(var) * z[d];
        const double delta = u[d] - uid;
        for (int64_t j = 0; j < nnz; j++) yhat[j] += delta * prod[(size_t)j * K + d];
      }
    }
    free(Q);
    free(z);
    free(yhat);
    free(prod);
  }
}

void orc_at_mul_bt_col(int nr, const orc_csr* Ft, int64_t f, const double* Z, double* y) {
  for (int d = 0; d < nr; d++) y[d] = 0.0;
  for (int64_t e = Ft->row_ptr[f]; e < Ft->row_ptr[f + 1]; e++) {
    const double v = Ft->vals ? Ft->vals[e] : 1.0;
    const double* z = Z + (size_t)Ft->cols[e] * nr;
    for (int d = 0; d < nr; d++) y[d] += v * z[d];
  }
}

void orc_add_acol_mul_bt(int nr, const orc_csr* Ft, int64_t f, const double* b, double* Z) {
  for (int64_t e = Ft->row_ptr[f]; e < Ft->row_ptr[f + 1]; e++) {
    const double v = Ft->vals ? Ft->vals[e] : 1.0;
    double* z = Z + (size_t)Ft->cols[e] * nr;
    for (int d = 0; d < nr; d++) z[d] += v * b[d];
  }
}

void orc_beta_sweep_uni(int K, int64_t N, const orc_csr* Ft, const double* colsq, const double* lambda,
                        const double* lambda_beta, const double* mu, const double* U, const double* Uhat,
                        double* beta, uint64_t seed, uint32_t it, int mode) {
  const int64_t nfeat = Ft->nrow;
  double* Z = (double*)malloc(sizeof(double) * (size_t)K * (size_t)(N > 0 ? N : 1));
  double* zx = (double*)malloc(sizeof(double) * K);
  double* delta = (double*)malloc(sizeof(double) * K);
  /* Z = U - mu - Uhat (macauoneprior.cpp:227-232) */
  for (int64_t i = 0; i < N; i++)
    for (int d = 0; d < K; d++) Z[(size_t)i * K + d] = U[(size_t)i * K + d] - mu[d] - Uhat[(size_t)i * K + d];
  /* features in order; the latents are independent chains, so all K are
   * advanced together (the reference runs them in blocks of 4, :220-251) */
  for (int64_t f = 0; f < nfeat; f++) {
    orc_at_mul_bt_col(K, Ft, f, Z, zx);
    for (int d = 0; d < K; d++) {
      const double A = lambda_beta[d] + lambda[d] * colsq[f];
      const double B = lambda[d] * (zx[d] + beta[(size_t)f * K + d] * colsq[f]);
      const double Ainv = 1.0 / A;
      const double bnew =
          B * Ainv + sqrt(Ainv) * orc_normal(seed, (uint32_t)f, (uint32_t)d, it, ORC_TAG(ORC_TAG_BETA_ONE, mode));
      delta[d] = beta[(size_t)f * K + d] - bnew;
      beta[(size_t)f * K + d] = bnew;
    }
    orc_add_acol_mul_bt(K, Ft, f, delta, Z);
  }
  free(Z);
  free(zx);
  free(delta);
}

void orc_lambda_beta_uni(int K, int64_t nfeat, const double* beta, uint64_t seed, uint32_t it, int mode,
                         double* lambda_beta) {
  for (int d = 0; d < K; d++) {
    double ss = 0.0;
    for (int64_t f = 0; f < nfeat; f++) ss += beta[(size_t)f * K + d] * beta[(size_t)f * K + d];
    const double a = 0.1 + (double)nfeat / 2.0, b = 0.1 + ss / 2.0;
    lambda_beta[d] = orc_gamma(seed, a, (uint32_t)d, it, ORC_TAG(ORC_TAG_LAMBDA_BETA, mode)) / b;
  }
}

int orc_macau_one_update_prior(int K, int mode, const double* U, int64_t N, const orc_csr* F, const orc_csr* Ft,
                               const double* colsq, double* mu, double* lambda, double* beta, double* Uhat,
                               double* lambda_beta, uint64_t seed, uint32_t it) {
  const size_t KK = (size_t)K * K;
  double* Ud = (double*)malloc(sizeof(double) * (size_t)K * N);
  double* WI = (double*)calloc(KK, sizeof(double));
  double* mu0 = (double*)calloc(K, sizeof(double));
  double* Lam = (double*)malloc(sizeof(double) * KK);
  int rc = 0;
  /* sample_mu_lambda (:198-213): the Normal-Wishart draw, diagonal kept */
  for (size_t q = 0; q < (size_t)K * N; q++) Ud[q] = U[q] - Uhat[q];
  for (int i = 0; i < K; i++) WI[i + (size_t)i * K] = 1.0;
  rc = orc_cond_normal_wishart(K, N, Ud, mu0, 2.0, WI, (double)K, seed, it, mode, mu, Lam);
  if (rc) goto out;
  for (int d = 0; d < K; d++) lambda[d] = Lam[d + (size_t)d * K];
  orc_beta_sweep_uni(K, N, Ft, colsq, lambda, lambda_beta, mu, U, Uhat, beta, seed, it, mode);  /* :215-252 */
  orc_csr_A_mul_Bn(K, F, beta, Uhat);                                                          /* compute_uhat */
  orc_lambda_beta_uni(K, F->ncol, beta, seed, it, mode, lambda_beta);                           /* :263-289 */
out:
  free(Ud); free(WI); free(mu0); free(Lam);
  return rc;
}

/* ------------------------------------------------------------------ */
/* probit noise                                                             */
/* ------------------------------------------------------------------ */
double orc_norm_cdf(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }

/* Wichura, "Algorithm AS 241: the percentage points of the normal
 * distribution", Applied Statistics 37 (1988) 477-484, PPND16: rational
 * minimax approximations in q = p - 1/2 (|q| <= 0.425) and in
 * r = sqrt(-log(min(p, 1 - p))) (r <= 5, r > 5); about 1e-16 relative. */
static const double AS241_A[8] = {3.3871328727963666080e0, 1.3314166789178437745e+2, 1.9715909503065514427e+3,
                                  1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4,
                                  3.3430575583588128105e+4, 2.5090809287301226727e+3};
static const double AS241_B[8] = {1.0, 4.2313330701600911252e+1, 6.8718700749205790830e+2, 5.3941960214247511077e+3,
                                  2.1213794301586595867e+4, 3.9307895800092710610e+4, 2.8729085735721942674e+4,
                                  5.2264952788528545610e+3};
static const double AS241_C[8] = {1.42343711074968357734e0, 4.63033784615654529590e0, 5.76949722146069140550e0,
                                  3.64784832476320460504e0, 1.27045825245236838258e0, 2.41780725177450611770e-1,
                                  2.27238449892691845833e-2, 7.74545014278341407640e-4};
static const double AS241_D[8] = {1.0, 2.05319162663775882187e0, 1.67638483018380384940e0, 6.89767334985100004550e-1,
                                  1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4,
                                  1.05075007164441684324e-9};
static const double AS241_E[8] = {6.65790464350110377720e0, 5.46378491116411436990e0, 1.78482653991729133580e0,
                                  2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3,
                                  2.71155556874348757815e-5, 2.01033439929228813265e-7};
static const double AS241_F[8] = {1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2,
                                  7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7,
                                  2.04426310338993978564e-15};

static double horner8(const double* c, double x) {
  double v = c[7];
  for (int i = 6; i >= 0; i--) v = v * x + c[i];
  return v;
}

double orc_inv_norm_cdf(double p) {
  if (!(p > 0.0)) return -INFINITY;
  if (!(p < 1.0)) return INFINITY;
  const double q = p - 0.5;
  if (fabs(q) <= 0.425) {
    const double r = 0.180625 - q * q;
    return q * horner8(AS241_A, r) / horner8(AS241_B, r);
  }
  double r = sqrt(-log(q < 0.0 ? p : 1.0 - p));
  double x;
  if (r <= 5.0) {
    r -= 1.6;
    x = horner8(AS241_C, r) / horner8(AS241_D, r);
  } else {
    r -= 5.0;
    x = horner8(AS241_E, r) / horner8(AS241_F, r);
  }
  return q < 0.0 ? -x : x;
}

/* (k + 0.5) / 2^53 for the 53-bit integer k of two Philox words */
static double ud53(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 6) << 27) | (uint64_t)(b >> 5);
  return ((double)k + 0.5) * (1.0 / 9007199254740992.0);
}

double orc_truncnorm0(double mean, uint64_t seed, uint32_t c0, uint32_t it, uint32_t tag) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const double abar = -mean;  /* (low_cut - mean) / std with low_cut 0, std 1 */
  uint32_t w[4];
  if (abar > 3.0) {
    /* rand_truncnorm_rej (truncnorm.cpp:14-24).  Its acceptance test
     * v <= xbar / abar holds for every draw (xbar >= abar), so the first
     * proposal is returned: xbar has the law P(xbar > x) = exp(-(x^2 -
     * abar^2) / 2), as in the reference */
    for (uint32_t t = 0; t < 64; t++) {
      const uint32_t ctr[4] = {c0, t, it, tag};
      orc_philox4x32_10(ctr, key, w);
      const double u = ud53(w[0], w[1]), v = ud53(w[2], w[3]);
      const double xbar = sqrt(abar * abar - 2.0 * log(1.0 - u));
      if (v <= xbar / abar) return xbar + mean;
    }
    return abar + mean;
  }
  /* rand_truncnorm_icdf (truncnorm.cpp:9-12): U(Phi(abar), 1) through Phi^{-1} */
  const uint32_t ctr[4] = {c0, 0u, it, tag};
  orc_philox4x32_10(ctr, key, w);
  const double lo = orc_norm_cdf(abar);
  return orc_inv_norm_cdf(lo + (1.0 - lo) * ud53(w[0], w[1])) + mean;
}

void orc_probit_z(int K, const orc_rows* rows, const double* const* others, const double* U, uint64_t seed,
                  uint32_t it, int mode, double* z) {
  const int no = rows->nother;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t n = 0; n < rows->nrows; n++) {
    const double* u = U + (size_t)n * K;
    for (int64_t e = rows->row_ptr[n]; e < rows->row_ptr[n + 1]; e++) {
      double dot = 0.0;
      for (int k = 0; k < K; k++) {
        double c = others[0][(size_t)rows->idx[e * no] * K + k];
        for (int t = 1; t < no; t++) c *= others[t][(size_t)rows->idx[e * no + t] * K + k];
        dot += c * u[k];
      }
      const double y = 2.0 * rows->vals[e] - 1.0;
      z[e] = y * orc_truncnorm0(y * dot, seed, (uint32_t)e, it, ORC_TAG(ORC_TAG_PROBIT, mode));
    }
  }
}

static const double* auc_key;
static int auc_cmp(const void* a, const void* b) {
  const int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
  if (auc_key[i] < auc_key[j]) return -1;
  if (auc_key[i] > auc_key[j]) return 1;
  return i < j ? -1 : (i > j ? 1 : 0);
}

double orc_auc(int64_t n, const double* pred, const double* test) {
  if (n == 0) return NAN;
  int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t i = 0; i < n; i++) perm[i] = i;
  auc_key = pred;
  qsort(perm, (size_t)n, sizeof(int64_t), auc_cmp);
  double NP = 0.0;
  for (int64_t i = 0; i < n; i++) NP += test[i];
  const double NN = (double)n - NP;
  /* stack_x = running positives, stack_y = running negatives along the sort */
  double sx = test[perm[0]], sy = 1.0 - sx, auc = 0.0;
  for (int64_t i = 1; i < n; i++) {
    const double x1 = sx + test[perm[i]], y1 = sy + 1.0 - test[perm[i]];
    auc += (x1 - sx) * y1;
    sx = x1;
    sy = y1;
  }
  free(perm);
  return auc / (NP * NN);
}

void orc_eval_probit(int K, int64_t ntest, const int32_t* test_idx, const double* test_vals, const double* const* U,
                     int n, double* pred, double* pred_var, double* pred1, double* auc1, double* auc_avg) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < ntest; t++) {
    const int32_t* ix = test_idx + t * 2;
    double d = 0.0;
    for (int k = 0; k < K; k++) d += U[0][(size_t)ix[0] * K + k] * U[1][(size_t)ix[1] * K + k];
    const double p = orc_norm_cdf(d);
    double pa;
    if (n == 0) {
      pa = p;
    } else {
      const double delta = p - pred[t];
      pa = pred[t] + delta / (n + 1);
      pred_var[t] += delta * (p - pa);
    }
    pred1[t] = p;
    pred[t] = pa;
  }
  *auc1 = orc_auc(ntest, pred1, test_vals);
  *auc_avg = orc_auc(ntest, pred, test_vals);
}

/* ------------------------------------------------------------------ */
/* a4 Gibbs driver (MacauX::init / run, macau.cpp:72-86, 94-148)            */
/* ------------------------------------------------------------------ */
int orc_run(orc_run_cfg* c) {
  const int K = c->K, M = c->nmodes;
  if (M < 2 || M > ORC_MAX_MODES) return -5;
  const size_t KK = (size_t)K * K;
  double* U[ORC_MAX_MODES] = {0};
  double* mu[ORC_MAX_MODES] = {0};
  double* Lambda[ORC_MAX_MODES] = {0};
  double* beta[ORC_MAX_MODES] = {0};
  double* Uhat[ORC_MAX_MODES] = {0};
  double* murows[ORC_MAX_MODES] = {0};
  double* FtF[ORC_MAX_MODES] = {0};
  double* lam_uni[ORC_MAX_MODES] = {0};   /* MacauOnePrior: diagonal prior precision */
  double* lb_uni[ORC_MAX_MODES] = {0};    /* MacauOnePrior: per-latent lambda_beta */
  double* colsq[ORC_MAX_MODES] = {0};     /* MacauOnePrior: squared column norms of F */
  double lambda_beta[ORC_MAX_MODES];
  double* WI = (double*)calloc(KK, sizeof(double));
  int rc = 0;
  for (int i = 0; i < K; i++) WI[i + i * K] = 1.0;
  for (int m = 0; m < M; m++) {
    U[m] = (double*)calloc((size_t)K * c->dims[m], sizeof(double)); /* init(): setZero */
    /* BPMFPrior::init / MacauPrior::init (latentprior.cpp:84-99, 181-216) */
    mu[m] = (double*)calloc(K, sizeof(double));
    Lambda[m] = (double*)calloc(KK, sizeof(double));
    for (int i = 0; i < K; i++) Lambda[m][i + i * K] = 10.0;
    if (c->macau[m]) {
      beta[m] = (double*)calloc((size_t)K * c->F[m].ncol, sizeof(double));
      Uhat[m] = (double*)calloc((size_t)K * c->dims[m], sizeof(double));
      murows[m] = (double*)malloc(sizeof(double) * (size_t)K * c->dims[m]);
      lambda_beta[m] = c->lambda_beta0;
      if (c->uni[m]) {  /* MacauOnePrior::init (macauoneprior.cpp:16-38), setLambdaBeta */
        lam_uni[m] = (double*)malloc(sizeof(double) * K);
        lb_uni[m] = (double*)malloc(sizeof(double) * K);
        colsq[m] = (double*)calloc((size_t)c->F[m].ncol + 1, sizeof(double));
        for (int k = 0; k < K; k++) {
          lam_uni[m][k] = 10.0;
          lb_uni[m][k] = c->lambda_beta0;
        }
        for (int64_t f = 0; f < c->F[m].ncol; f++)
          for (int64_t e = c->Ft[m].row_ptr[f]; e < c->Ft[m].row_ptr[f + 1]; e++)
            colsq[m][f] += c->Ft[m].vals ? c->Ft[m].vals[e] * c->Ft[m].vals[e] : 1.0;
      } else if (c->use_ftf[m]) {  /* MacauPrior::init: At_mul_A(FtF, *F) (latentprior.cpp:200-203) */
        FtF[m] = (double*)malloc(sizeof(double) * (size_t)c->F[m].ncol * (size_t)c->F[m].ncol);
        orc_at_mul_a(&c->F[m], &c->Ft[m], FtF[m]);
      }
    }
  }
  /* noise.init (noisemodels.cpp:14-53) */
  double alpha = c->precision, alpha_max = 0.0, var_total = 1.0;
  if (c->noise_adaptive) {
    double se = 0.0;
    const orc_rows* r0 = &c->train[0];
    for (int64_t e = 0; e < r0->row_ptr[r0->nrows]; e++) se += (r0->vals[e] - c->mean) * (r0->vals[e] - c->mean);
    var_total = se / (double)c->nnz;
    if (var_total <= 0.0 || isnan(var_total)) var_total = 1.0;
    alpha = (c->sn_init + 1.0) / var_total;
    alpha_max = (c->sn_max + 1.0) / var_total;
  }
  for (int64_t t = 0; t < c->ntest; t++) { c->pred[t] = 0.0; c->pred_var[t] = 0.0; }
  c->rmse = NAN;
  c->rmse1 = NAN;
  c->auc = NAN;
  c->auc1 = NAN;
  /* probit: the samplers see precision 1 and the augmented z as the
   * observations, with no mean (latentprior.cpp:387-414) */
  double* zbuf = NULL;
  double* pred1 = NULL;
  if (c->noise_probit) {
    if (M != 2) { rc = -3; goto done; }  /* "Unimplemented: sample_latents" (latentprior.cpp:112-116) */
    for (int m = 0; m < M; m++)
      if (c->macau[m] && c->uni[m]) { rc = -4; goto done; }  /* "Not implemented!" (macauoneprior.cpp:260-264) */
    int64_t mx = 0;
    for (int m = 0; m < M; m++) mx = c->train[m].row_ptr[c->train[m].nrows] > mx ? c->train[m].row_ptr[c->train[m].nrows] : mx;
    zbuf = (double*)malloc(sizeof(double) * ((size_t)mx + 1));
    pred1 = (double*)malloc(sizeof(double) * ((size_t)c->ntest + 1));
    alpha = 1.0;
  }
  int total = c->burnin + c->nsamples;
  if (c->max_iters > 0 && c->max_iters < total) total = c->max_iters;
  for (int i = 0; i < total; i++) {
#if ORC_MUTANT == 5
    for (int m = 0; m < M; m++) {
      double* mu0 = (double*)calloc(K, sizeof(double));
      if (!c->macau[m])
        rc = orc_cond_normal_wishart(K, c->dims[m], U[m], mu0, 2.0, WI, (double)K, c->seed, (uint32_t)i, m, mu[m],
                                     Lambda[m]);
      free(mu0);
    }
#endif
    for (int m = 0; m < M; m++) {
      const double* others[ORC_MAX_MODES];
      int t = 0;
      for (int q = 0; q < M; q++)
        if (q != m) others[t++] = U[q];
      if (c->macau[m]) {
        for (int64_t n = 0; n < c->dims[m]; n++)
          for (int k = 0; k < K; k++) murows[m][(size_t)n * K + k] = mu[m][k] + Uhat[m][(size_t)n * K + k];
      }
      int64_t f;
      if (c->macau[m] && c->uni[m]) {
        orc_sample_latents_uni(K, &c->train[m], c->mean, others, alpha, lam_uni[m], murows[m], c->seed, (uint32_t)i,
                               m, U[m]);
        f = 0;
      } else if (c->noise_probit) {
        orc_probit_z(K, &c->train[m], others, U[m], c->seed, (uint32_t)i, m, zbuf);
        orc_rows zr = c->train[m];
        zr.vals = zbuf;
        f = orc_sample_latents(K, &zr, 0.0, others, 1.0, mu[m], Lambda[m], c->macau[m] ? murows[m] : NULL, c->seed,
                               (uint32_t)i, m, U[m]);
      } else {
        f = orc_sample_latents(K, &c->train[m], c->mean, others, alpha, mu[m], Lambda[m],
                               c->macau[m] ? murows[m] : NULL, c->seed, (uint32_t)i, m, U[m]);
      }
      if (f) { rc = -1; goto done; }
#if ORC_MUTANT == 7 /* planted mistake (tests only): noise precision updated between the modes */
      if (m == 0 && c->noise_adaptive && !c->noise_probit) {
        const double* Us[ORC_MAX_MODES];
        for (int q = 0; q < M; q++) Us[q] = U[q];
        alpha = orc_noise_alpha(orc_train_sse(K, M, &c->train[0], c->mean, Us), c->nnz, var_total, c->is_tensor,
                                alpha_max, c->seed, (uint32_t)i);
      }
#endif
    }
    for (int m = 0; m < M; m++) {
#if ORC_MUTANT == 5 /* planted mistake (tests only): the priors updated before the modes are sampled */
      if (!c->macau[m]) continue;
#endif
      if (c->macau[m] && c->uni[m]) {
        rc = orc_macau_one_update_prior(K, m, U[m], c->dims[m], &c->F[m], &c->Ft[m], colsq[m], mu[m], lam_uni[m],
                                        beta[m], Uhat[m], lb_uni[m], c->seed, (uint32_t)i);
      } else if (c->macau[m]) {
        int cgit = 0;
        rc = orc_macau_update_prior(K, m, U[m], c->dims[m], &c->F[m], &c->Ft[m], FtF[m], WI, K, beta[m], Uhat[m],
                                    &lambda_beta[m], c->tol, mu[m], Lambda[m], c->seed, (uint32_t)i, &cgit);
        c->cg_iters_last = cgit;
      } else {
        double* mu0 = (double*)calloc(K, sizeof(double));
        rc = orc_cond_normal_wishart(K, c->dims[m], U[m], mu0, 2.0, WI, (double)K, c->seed, (uint32_t)i, m,
                                     mu[m], Lambda[m]);
        free(mu0);
      }
      if (rc) { rc = -2; goto done; }
    }
    if (c->noise_adaptive && !c->noise_probit && ORC_MUTANT != 7) {
      const double* Us[ORC_MAX_MODES];
      for (int m = 0; m < M; m++) Us[m] = U[m];
      double sse = orc_train_sse(K, M, &c->train[0], c->mean, Us);
      alpha = orc_noise_alpha(sse, c->nnz, var_total, c->is_tensor, alpha_max, c->seed, (uint32_t)i);
    }
    {
      const double* Us[ORC_MAX_MODES];
      for (int m = 0; m < M; m++) Us[m] = U[m];
      if (c->noise_probit)
        orc_eval_probit(K, c->ntest, c->test_idx, c->test_vals, Us, i < c->burnin ? 0 : i - c->burnin, c->pred,
                        c->pred_var, pred1, &c->auc1, &c->auc);
      else
        orc_eval_rmse(K, M, c->ntest, c->test_idx, c->test_vals, c->mean, Us, i < c->burnin ? 0 : i - c->burnin,
                      c->pred, c->pred_var, &c->rmse1, &c->rmse);
    }
    if (c->alpha_trace) c->alpha_trace[i] = alpha;
    if (c->rmse_trace) c->rmse_trace[i] = c->noise_probit ? c->auc : c->rmse;
  }
  c->alpha = alpha;
  for (int m = 0; m < M; m++) {
    if (c->U_out[m]) memcpy(c->U_out[m], U[m], sizeof(double) * (size_t)K * c->dims[m]);
    if (c->mu_out[m]) memcpy(c->mu_out[m], mu[m], sizeof(double) * K);
    if (c->Lambda_out[m]) memcpy(c->Lambda_out[m], Lambda[m], sizeof(double) * KK);
    if (c->lambda_beta_out && c->macau[m] && c->uni[m]) {
      double s = 0.0;
      for (int k = 0; k < K; k++) s += lb_uni[m][k];
      c->lambda_beta_out[m] = s / K;  /* getLinkLambda: lambda_beta.mean() (macauoneprior.h:45) */
    } else if (c->lambda_beta_out && c->macau[m]) {
      c->lambda_beta_out[m] = lambda_beta[m];
    }
  }
done:
  for (int m = 0; m < M; m++) {
    free(U[m]); free(mu[m]); free(Lambda[m]); free(beta[m]); free(Uhat[m]); free(murows[m]); free(FtF[m]);
    free(lam_uni[m]); free(lb_uni[m]); free(colsq[m]);
  }
  free(WI);
  free(zbuf);
  free(pred1);
  return rc;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
