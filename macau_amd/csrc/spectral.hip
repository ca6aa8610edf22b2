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

constexpr int kSmRows = 128;        // output rows per work-group (32 per wave)
constexpr int kSmK = 32;            // right-hand sides per work-group
constexpr int kSmT = 32;            // inner-dimension slab staged in LDS
// LDS strides chosen so each 64-lane operand read (16 rows x 4 inner
// indices, or 4 inner rows x 16 columns) takes the minimum two bank passes
constexpr int kMsLd = kSmT + 4;
constexpr int kXsLd = kSmK + 16;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// part[blockIdx.y][r][k] = sum over t in chunk blockIdx.y of M[t + r n] X[t][k]
// (M column-major n x n, so the sum runs down column r of M: out = M^T X).
// Work-group (x, y, z): rows r0 .. r0 + 127, inner chunk y, columns
// k0 .. k0 + 31.  Wave w owns rows r0 + 32 w .. + 31 as 2 x 2 FP64 MFMA
// tiles of 16 x 16 (v_mfma_f64_16x16x4f64: lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16] and D[4 v + l / 16][l % 16], v = 0..3).  The next
// slab's global loads are issued before the current slab is consumed.
__global__ void __launch_bounds__(256, 3) k_spec_mm(const double* __restrict__ M, int n, const double* __restrict__ X,
                                                 int nrhs, int tchunk, double* __restrict__ part) {
  __shared__ __align__(16) double Ms[kSmRows][kMsLd];  // [row][inner]
  __shared__ __align__(16) double Xs[kSmT][kXsLd];     // [inner][column]
  const int r0 = blockIdx.x * kSmRows;
  const int k0 = blockIdx.z * kSmK;
  const int tb = blockIdx.y * tchunk;
  const int te = min(n, tb + tchunk);
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
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
  f64x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
  if (tb < te) load(tb);
  for (int t0 = tb; t0 < te; t0 += kSmT) {
    __syncthreads();  // the previous slab has been consumed
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = tid + 256 * u;
      Ms[q >> 5][q & 31] = mv[u];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u;
      Xs[q >> 5][q & 31] = xv[u];
    }
    __syncthreads();
    if (t0 + kSmT < te) load(t0 + kSmT);
#pragma unroll
    for (int st = 0; st < kSmT / 4; st++) {
      const int tt = 4 * st + lk;
      const double a0 = Ms[32 * w + li][tt], a1 = Ms[32 * w + 16 + li][tt];
      const double b0 = Xs[tt][li], b1 = Xs[tt][16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  double* out = part + (size_t)blockIdx.y * n * nrhs;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int k = k0 + 16 * b + li;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = r0 + 32 * w + 16 * a + 4 * v + lk;
        if (r < n && k < nrhs) out[(size_t)r * nrhs + k] = acc[a][b][v];
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
  // enough inner chunks for one resident round of work-groups (three per
  // CU on 256 CUs: 768), each at least 128 deep; the slab loop handles a
  // chunk depth that is not a multiple of the slab
  int S = (768 + row_tiles * k_tiles - 1) / (row_tiles * k_tiles);
  const int smax = (n + 127) / 128;
  if (S > smax) S = smax;
  if (S < 1) S = 1;
  const int tchunk = (n + S - 1) / S;
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
