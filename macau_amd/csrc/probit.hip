// probit.hip -- data augmentation of the probit noise model.
//
// ProbitNoise (noisemodels.h:88-103) samples a row u with precision 1 and
// the observations replaced by latent Gaussians
//   z_e = y rand_truncnorm(y u.v_e, 1, 0),   y = 2 value - 1
// drawn from the row's value before its update (sample_latent_blas_probit,
// latentprior.cpp:387-414).  k_probit_z draws z for every observation of one
// mode; the samplers then read z in place of y - mean, with alpha = 1.
//
// rand_truncnorm(mean, 1, 0) (truncnorm.cpp:9-37): with the cut
// abar = -mean, inverse-CDF sampling of U(Phi(abar), 1) for abar <= 3 and
// the reference's tail sampler above.  That sampler's acceptance test
// v <= xbar / abar holds for every proposal xbar = sqrt(abar^2 - 2 log(1-u))
// >= abar, so tail draws follow the proposal law, as in the reference.  Phi^{-1} starts from Abramowitz
// & Stegun 26.2.23 (|error| < 4.5e-4) and takes three Halley steps on
// Phi(y) = min(p, 1 - p).  Uniforms have 53 bits, from Philox block
// (observation, attempt, iteration, tag(PROBIT, mode)).
//
// Bytes per observation: idx 4 + row 4 + value 8 + z 4, plus the two factor
// rows (2 KP x 4, mostly L2 hits); FP64 work: three erfc/exp pairs, a log
// and a sqrt.
#include <math.h>

#include "common.h"
#include "kernels.h"

namespace mgpu {

__device__ __forceinline__ double norm_cdf_d(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }

__device__ double inv_norm_cdf_d(double p) {
  if (!(p > 0.0)) return -INFINITY;
  if (!(p < 1.0)) return INFINITY;
  const double q = p < 0.5 ? p : 1.0 - p;
  const double t = sqrt(-2.0 * log(q));
  double y = -(t - (2.515517 + t * (0.802853 + t * 0.010328)) / (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308))));
  for (int k = 0; k < 3 && q >= 2.2250738585072014e-308; k++) {
    const double e = norm_cdf_d(y) - q;
    const double u = e * 2.5066282746310002 * exp(0.5 * y * y);
    y -= u / (1.0 + 0.5 * y * u);
  }
  return p < 0.5 ? y : -y;
}

__device__ __forceinline__ double ud53(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 6) << 27) | (uint64_t)(b >> 5);
  return ((double)k + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ double truncnorm0(double mean, uint64_t seed, uint32_t c0, uint32_t it, uint32_t tag) {
  const double abar = -mean;
  uint32_t w[4];
  if (abar > 3.0) {
    for (uint32_t t = 0; t < 64; t++) {
      philox_at(seed, c0, t, it, tag, w);
      const double u = ud53(w[0], w[1]), v = ud53(w[2], w[3]);
      const double xbar = sqrt(abar * abar - 2.0 * log(1.0 - u));
      if (v <= xbar / abar) return xbar + mean;
    }
    return abar + mean;
  }
  philox_at(seed, c0, 0u, it, tag, w);
  const double lo = norm_cdf_d(abar);
  return inv_norm_cdf_d(lo + (1.0 - lo) * ud53(w[0], w[1])) + mean;
}

constexpr int kProbitThreads = 256;

template <int KP>
__global__ void __launch_bounds__(kProbitThreads) k_probit_z(ProbitZArgs a) {
  const uint32_t it = (uint32_t)*a.iter;
  const uint32_t tag = make_tag(TAG_PROBIT, (uint32_t)a.mode);
  for (long long e = (long long)blockIdx.x * kProbitThreads + threadIdx.x; e < a.nnz;
       e += (long long)gridDim.x * kProbitThreads) {
    const float4* u = reinterpret_cast<const float4*>(a.U + (size_t)a.row_of[e] * KP);
    const float4* v = reinterpret_cast<const float4*>(a.V + (size_t)a.idx[e] * KP);
    float d0 = 0.f, d1 = 0.f;
#pragma unroll
    for (int q = 0; q < KP / 4; q++) {
      const float4 x = u[q], w = v[q];
      d0 = fmaf(x.x, w.x, d0);
      d1 = fmaf(x.y, w.y, d1);
      d0 = fmaf(x.z, w.z, d0);
      d1 = fmaf(x.w, w.w, d1);
    }
    const double y = 2.0 * a.y[e] - 1.0;
    a.z[e] = (float)(y * truncnorm0(y * (double)(d0 + d1), a.seed, (uint32_t)e, it, tag));
  }
}

void launch_probit_z(const ProbitZArgs& a, hipStream_t s) {
  if (a.nnz <= 0) return;
  long long b = (a.nnz + kProbitThreads - 1) / kProbitThreads;
  if (b > 8192) b = 8192;
  const dim3 g((unsigned)b), t(kProbitThreads);
  switch (a.KP) {
    case 8: hipLaunchKernelGGL(k_probit_z<8>, g, t, 0, s, a); break;
    case 16: hipLaunchKernelGGL(k_probit_z<16>, g, t, 0, s, a); break;
    case 32: hipLaunchKernelGGL(k_probit_z<32>, g, t, 0, s, a); break;
    case 64: hipLaunchKernelGGL(k_probit_z<64>, g, t, 0, s, a); break;
    default: hipLaunchKernelGGL(k_probit_z<128>, g, t, 0, s, a); break;
  }
}

// Truncated-normal probes for the parity tests: out[j] = rand_truncnorm(
// mean[j], 1, 0) keyed by (c0 + j, it, tag); inv[j] = Phi^{-1}(p[j]).
__global__ void k_truncnorm_probe(int n, const double* mean, const double* p, uint64_t seed, uint32_t c0,
                                  uint32_t it, uint32_t tag, double* out, double* inv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  out[j] = truncnorm0(mean[j], seed, c0 + (uint32_t)j, it, tag);
  inv[j] = inv_norm_cdf_d(p[j]);
}
void launch_truncnorm_probe(int n, const double* mean, const double* p, uint64_t seed, uint32_t c0, uint32_t it,
                            uint32_t tag, double* out, double* inv, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_truncnorm_probe, dim3((n + 255) / 256), dim3(256), 0, s, n, mean, p, seed, c0, it,
                                tag, out, inv);
}

}  // namespace mgpu
