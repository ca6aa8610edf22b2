// spectral.hip -- the F^T F branch of MacauPrior::sample_beta
// (latentprior.cpp:280-289) through one eigendecomposition of F^T F.
//
// The reference solves (F^T F + lambda_beta I) X = B with a fresh Cholesky
// factorization every Gibbs iteration (chol_decomp, chol_solve_t,
// chol.cpp:13-53).  F^T F is fixed for the whole chain and only lambda_beta
// changes, so F^T F = Q diag(d) Q^T, computed once at init (rocSOLVER dsyevd,
// FP64), turns every later solve into
//     Y = Q^T B,   Y[j][:] /= (d_j + lambda_beta),   X = Q Y,
// two n x n by n x nrhs FP64 products that stream Q from HBM once each, in
// place of n / 32 dependent factor and solve steps.  Q is kept twice (Q and
// Q^T, both column-major) so both products read M's columns contiguously.
// A non-positive d_j + lambda_beta sets the same fail flag as a non-positive
// Cholesky pivot.
#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <rocsolver/rocsolver.h>
#include <stdexcept>
#include <string>

#include "common.h"
#include "kernels.h"

namespace mgpu {

constexpr int kSmRows = 128;        // output rows per work-group (4 per thread)
constexpr int kSmK = 32;            // right-hand sides per work-group (4 per thread)
constexpr int kSmT = 32;            // inner-dimension slab staged in LDS
constexpr int kSmLd = kSmRows + 2;  // LDS row stride of the M slab (16-byte aligned rows)

// part[blockIdx.y][r][k] = sum over t in chunk blockIdx.y of M[t + r n] X[t][k]
// (M column-major n x n, so the sum runs down column r of M: out = M^T X).
// Work-group (x, y, z): rows r0 .. r0 + 127, inner chunk y, columns
// k0 .. k0 + 31; thread (rg, kg) keeps a 4 x 4 register tile.  The next
// slab's global loads are issued before the current slab is consumed.
__global__ void __launch_bounds__(256) k_spec_mm(const double* __restrict__ M, int n, const double* __restrict__ X,
                                                 int nrhs, int tchunk, double* __restrict__ part) {
  __shared__ __align__(16) double Ms[kSmT][kSmLd];
  __shared__ __align__(16) double Xs[kSmT][kSmK];
  const int r0 = blockIdx.x * kSmRows;
  const int k0 = blockIdx.z * kSmK;
  const int tb = blockIdx.y * tchunk;
  const int te = min(n, tb + tchunk);
  const int tid = threadIdx.x;
  const int rg = tid & 31, kg = tid >> 5;
  double mv[16], xv[4];
  auto load = [&](int t0) {
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = tid + 256 * u;
      const int gt = t0 + (q & 31), gr = r0 + (q >> 5);
      mv[u] = (gt < te && gr < n) ? M[(size_t)gt + (size_t)gr * n] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u;
      const int gt = t0 + (q >> 5), gk = k0 + (q & 31);
      xv[u] = (gt < te && gk < nrhs) ? X[(size_t)gt * nrhs + gk] : 0.0;
    }
  };
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
  if (tb < te) load(tb);
  for (int t0 = tb; t0 < te; t0 += kSmT) {
    __syncthreads();  // the previous slab has been consumed
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = tid + 256 * u;
      Ms[q & 31][q >> 5] = mv[u];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u;
      Xs[q >> 5][q & 31] = xv[u];
    }
    __syncthreads();
    if (t0 + kSmT < te) load(t0 + kSmT);
#pragma unroll 8
    for (int t = 0; t < kSmT; t++) {
      const double2 m01 = *reinterpret_cast<const double2*>(&Ms[t][4 * rg]);
      const double2 m23 = *reinterpret_cast<const double2*>(&Ms[t][4 * rg + 2]);
      const double2 x01 = *reinterpret_cast<const double2*>(&Xs[t][4 * kg]);
      const double2 x23 = *reinterpret_cast<const double2*>(&Xs[t][4 * kg + 2]);
      const double m[4] = {m01.x, m01.y, m23.x, m23.y};
      const double x[4] = {x01.x, x01.y, x23.x, x23.y};
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = fma(m[a], x[b], acc[a][b]);
    }
  }
  double* out = part + (size_t)blockIdx.y * n * nrhs;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int r = r0 + 4 * rg + a;
    if (r >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int k = k0 + 4 * kg + b;
      if (k < nrhs) out[(size_t)r * nrhs + k] = acc[a][b];
    }
  }
}

// out[r][k] = sum_s part[s][r][k] in fixed order s = 0 .. S-1, divided by
// d_r + *reg when SCALE (a non-positive divisor sets *fail and is taken as 1)
template <bool SCALE>
__global__ void __launch_bounds__(256) k_spec_reduce(const double* __restrict__ part, int S, int n, int nrhs,
                                                     const double* __restrict__ d, const double* reg,
                                                     double* __restrict__ out, int* fail) {
  const size_t tot = (size_t)n * nrhs;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < S; p++) s += part[(size_t)p * tot + q];
    if (SCALE) {
      double den = d[q / nrhs] + *reg;
      if (!(den > 0.0)) {
        atomicMax(fail, kFailDenseChol);
        den = 1.0;
      }
      s /= den;
    }
    out[q] = s;
  }
}

// Qt = Q^T for column-major n x n matrices, 32 x 32 tiles through LDS
__global__ void __launch_bounds__(256) k_transpose(const double* __restrict__ Q, int n, double* __restrict__ Qt) {
  __shared__ double T[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int c = ty; c < 32; c += 8) {
    const int i = i0 + tx, j = j0 + c;
    T[c][tx] = (i < n && j < n) ? Q[(size_t)i + (size_t)j * n] : 0.0;
  }
  __syncthreads();
  for (int c = ty; c < 32; c += 8) {
    const int i = j0 + tx, j = i0 + c;  // Qt[i + j n] = Q[j + i n]
    if (i < n && j < n) Qt[(size_t)i + (size_t)j * n] = T[tx][c];
  }
}

SpectralPlan spectral_plan(int n, int nrhs) {
  SpectralPlan p;
  const int row_tiles = (n + kSmRows - 1) / kSmRows;
  const int k_tiles = (nrhs + kSmK - 1) / kSmK;
  // enough inner chunks for about 1024 work-groups, each at least 128 deep
  int S = (1024 + row_tiles * k_tiles - 1) / (row_tiles * k_tiles);
  const int smax = (n + 127) / 128;
  if (S > smax) S = smax;
  if (S < 1) S = 1;
  int tchunk = (n + S - 1) / S;
  tchunk = (tchunk + kSmT - 1) / kSmT * kSmT;
  p.S = (n + tchunk - 1) / tchunk;
  p.tchunk = tchunk;
  return p;
}

static int reduce_grid(size_t work) {
  size_t g = (work + 255) / 256;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

void launch_spectral_solve(const double* Q, const double* Qt, const double* d, int n, const double* reg,
                           const double* B, double* Y, double* part, const SpectralPlan& p, double* X, int nrhs,
                           int* fail, hipStream_t s) {
  const dim3 grid((n + kSmRows - 1) / kSmRows, p.S, (nrhs + kSmK - 1) / kSmK);
  const int rg = reduce_grid((size_t)n * nrhs);
  // Y = diag(d + reg)^{-1} Q^T B
  hipLaunchKernelGGL(k_spec_mm, grid, dim3(256), 0, s, Q, n, B, nrhs, p.tchunk, part);
  hipLaunchKernelGGL(k_spec_reduce<true>, dim3(rg), dim3(256), 0, s, (const double*)part, p.S, n, nrhs, d, reg, Y,
                     fail);
  // X = Q Y: column r of Q^T is row r of Q
  hipLaunchKernelGGL(k_spec_mm, grid, dim3(256), 0, s, Qt, n, (const double*)Y, nrhs, p.tchunk, part);
  hipLaunchKernelGGL(k_spec_reduce<false>, dim3(rg), dim3(256), 0, s, (const double*)part, p.S, n, nrhs, nullptr,
                     nullptr, X, fail);
}

// rocSOLVER and the rocBLAS it was built against, resolved from one loaded
// copy of the library: a process that already holds a rocSOLVER (a framework
// that bundles its own) gets that copy and its own rocBLAS, never a mix.
namespace {
struct SolverApi {
  decltype(&rocblas_create_handle) create = nullptr;
  decltype(&rocblas_destroy_handle) destroy = nullptr;
  decltype(&rocblas_set_stream) set_stream = nullptr;
  decltype(&rocsolver_dsyevd) dsyevd = nullptr;
  std::string error;
};

const SolverApi& solver_api() {
  static const SolverApi api = [] {
    SolverApi a;
    void* h = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("librocsolver.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      const char* e = dlerror();
      a.error = std::string("cannot load rocSOLVER: ") + (e ? e : "unknown error");
      return a;
    }
    a.create = reinterpret_cast<decltype(a.create)>(dlsym(h, "rocblas_create_handle"));
    a.destroy = reinterpret_cast<decltype(a.destroy)>(dlsym(h, "rocblas_destroy_handle"));
    a.set_stream = reinterpret_cast<decltype(a.set_stream)>(dlsym(h, "rocblas_set_stream"));
    a.dsyevd = reinterpret_cast<decltype(a.dsyevd)>(dlsym(h, "rocsolver_dsyevd"));
    if (!a.create || !a.destroy || !a.set_stream || !a.dsyevd) a.error = "rocSOLVER lacks rocsolver_dsyevd or rocBLAS";
    return a;
  }();
  return api;
}
}  // namespace

bool spectral_factor(double* A, int n, double* d, double* Qt, hipStream_t s, std::string* why) {
  const SolverApi& api = solver_api();
  if (!api.error.empty()) {
    if (why) *why = api.error;
    return false;
  }
  rocblas_handle h = nullptr;
  if (api.create(&h) != rocblas_status_success) {
    if (why) *why = "rocblas_create_handle failed";
    return false;
  }
  double* E = nullptr;
  int* info = nullptr;
  bool ok = api.set_stream(h, s) == rocblas_status_success && hipMalloc(&E, (size_t)n * sizeof(double)) == hipSuccess &&
            hipMalloc(&info, sizeof(int)) == hipSuccess;
  if (ok) ok = api.dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, A, n, d, E, info) == rocblas_status_success;
  int hinfo = -1;
  if (ok) ok = hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess && hinfo == 0;
  if (!ok && why) *why = "rocsolver_dsyevd failed (info " + std::to_string(hinfo) + ")";
  if (ok) {
    hipLaunchKernelGGL(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, s, (const double*)A, n, Qt);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
    if (!ok && why) *why = "transposing the eigenvectors failed";
  }
  api.destroy(h);
  if (E) (void)hipFree(E);
  if (info) (void)hipFree(info);
  return ok;
}

}  // namespace mgpu
