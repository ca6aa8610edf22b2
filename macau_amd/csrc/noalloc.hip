// noalloc.hip -- CondNormalWishart on caller-owned device memory, with no
// device allocation: its working memory is one file-scope device array.
//
// The same kernels as macau_cond_normal_wishart (k_hyper_stats, k_stats_reduce,
// k_nw_draw_w), launched asynchronously on the caller's stream.  Working
// memory (per GPU, held until the process ends):
//   partial   kNaBlocks x (K + K^2) FP64 statistics partials
//   Lambda_f  64 x 64 FP32 and Lmu_f 64 FP32 (the draw's sampler copies)
//   iter, fail
// Its host address is never used: the device address of each GPU's copy is
// taken once per device with hipGetSymbolAddress and used everywhere.
// Calls that overlap on one GPU share this memory and race.
#include <mutex>
#include <stdexcept>
#include <string>

#include "model.h"

namespace mgpu {

namespace {

constexpr int kNaMaxK = 64;     // the one-wave draw's limit
constexpr int kNaBlocks = 64;   // statistics work-groups (columns are split among them)
constexpr size_t kNaPartial = (size_t)kNaBlocks * (kNaMaxK + kNaMaxK * kNaMaxK) * sizeof(double);
constexpr size_t kNaLf = (size_t)kNaMaxK * kNaMaxK * sizeof(float);
constexpr size_t kNaLmu = (size_t)kNaMaxK * sizeof(float);
constexpr size_t kNaOffLf = kNaPartial;
constexpr size_t kNaOffLmu = kNaOffLf + kNaLf;
constexpr size_t kNaOffInts = kNaOffLmu + kNaLmu;
constexpr size_t kNaBytes = kNaOffInts + 256;  // 2,146,560 bytes

static_assert(kNaPartial % 256 == 0 && kNaLf % 256 == 0 && kNaLmu % 256 == 0, "sections stay 256-byte aligned");

}  // namespace

__device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's copy of `scratch` (taken once per device,
// with that device current)
void* scratch_on_current_device() {
  int dev = -1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

__global__ void k_na_begin(int* ints, int iter) {
  if (threadIdx.x == 0) {
    ints[0] = iter;  // read by the draw through NWArgs::iter
    ints[1] = 0;     // failure flag: cleared on every call
  }
}

__global__ void k_na_status(const int* ints, int* status) {
  if (threadIdx.x == 0) *status = ints[1];
}

}  // namespace

void cond_normal_wishart_noalloc(int K, int N, const double* U, const double* mu0, double kappa, const double* T,
                                 double nu, uint64_t seed, int iter, int mode, double* mu_out, double* Lambda_out,
                                 int* status, hipStream_t s) {
  if (K < 1 || K > kNaMaxK)
    throw std::runtime_error("num_latent " + std::to_string(K) + " exceeds the allocation-free limit of " +
                             std::to_string(kNaMaxK) + " (range [1, 64])");
  if (N < 1) throw std::runtime_error("the allocation-free Normal-Wishart draw needs N >= 1 columns");
  if (!U || !mu0 || !T || !mu_out || !Lambda_out || !status)
    throw std::runtime_error("null device pointer");
  unsigned char* base = static_cast<unsigned char*>(scratch_on_current_device());
  double* partial = reinterpret_cast<double*>(base);
  int* ints = reinterpret_cast<int*>(base + kNaOffInts);
  int nblk = hyper_stats_blocks(N);
  if (nblk > kNaBlocks) nblk = kNaBlocks;
  hipLaunchKernelGGL(k_na_begin, dim3(1), dim3(64), 0, s, ints, iter);
  launch_hyper_stats_f64(U, K, N, K, partial, nblk, s);
  NWArgs a;
  a.K = K;
  a.KP = pad_latent(K);
  a.partial = partial;
  a.nblk = 1;  // reduced into the first partial row
  a.N = N;
  a.kappa = kappa;
  a.mu0 = mu0;
  a.WI = T;
  a.nu = nu;
  a.seed = seed;
  a.iter = ints;
  a.mode = mode;
  a.ws = nullptr;  // K <= 64: the draw's matrices live in LDS
  a.mu = mu_out;
  a.Lambda = Lambda_out;
  a.Lambda_f = reinterpret_cast<float*>(base + kNaOffLf);
  a.Lmu_f = reinterpret_cast<float*>(base + kNaOffLmu);
  a.fail = ints + 1;
  launch_nw_draw(a, s);
  hipLaunchKernelGGL(k_na_status, dim3(1), dim3(64), 0, s, ints, status);
  hip_check(hipGetLastError(), "allocation-free Normal-Wishart launch");
}

}  // namespace mgpu
