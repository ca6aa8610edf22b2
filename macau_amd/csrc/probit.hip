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
// >= abar, so tail draws follow the proposal law, as in the reference.
// Phi^{-1} is Wichura's AS 241 (about 1e-16 relative: one log and sqrt
// outside the central interval, two degree-7 polynomials and a division).
// Uniforms have 53 bits, from Philox block
// (observation, attempt, iteration, tag(PROBIT, mode)).
//
// Bytes per observation: idx 4 + row 4 + value 8 + z 4, plus the two factor
// rows (2 KP x 4, mostly L2 hits); FP64 work: one erfc for Phi(abar), a
// log and a sqrt, and the AS 241 polynomials.
#include <math.h>

#include "common.h"
#include "kernels.h"

namespace mgpu {

__device__ __forceinline__ double norm_cdf_d(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }

// Wichura, "Algorithm AS 241: the percentage points of the normal
// distribution", Applied Statistics 37 (1988) 477-484, PPND16.
struct AS241 {
  double c[8];
};
__constant__ AS241 kAS241A = {{3.3871328727963666080e0, 1.3314166789178437745e+2, 1.9715909503065514427e+3,
                                       1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4,
                                       3.3430575583588128105e+4, 2.5090809287301226727e+3}};
__constant__ AS241 kAS241B = {{1.0, 4.2313330701600911252e+1, 6.8718700749205790830e+2,
                                       5.3941960214247511077e+3, 2.1213794301586595867e+4, 3.9307895800092710610e+4,
                                       2.8729085735721942674e+4, 5.2264952788528545610e+3}};
__constant__ AS241 kAS241C = {{1.42343711074968357734e0, 4.63033784615654529590e0, 5.76949722146069140550e0,
                                       3.64784832476320460504e0, 1.27045825245236838258e0, 2.41780725177450611770e-1,
                                       2.27238449892691845833e-2, 7.74545014278341407640e-4}};
__constant__ AS241 kAS241D = {{1.0, 2.05319162663775882187e0, 1.67638483018380384940e0,
                                       6.89767334985100004550e-1, 1.48103976427480074590e-1, 1.51986665636164571966e-2,
                                       5.47593808499534494600e-4, 1.05075007164441684324e-9}};
__constant__ AS241 kAS241E = {{6.65790464350110377720e0, 5.46378491116411436990e0, 1.78482653991729133580e0,
                                       2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3,
                                       2.71155556874348757815e-5, 2.01033439929228813265e-7}};
__constant__ AS241 kAS241F = {{1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1,
                                       1.48753612908506148525e-2, 7.86869131145613259100e-4, 1.84631831751005468180e-5,
                                       1.42151175831644588870e-7, 2.04426310338993978564e-15}};

__device__ __forceinline__ double horner8(const AS241& c, double x) {
  double v = c.c[7];
#pragma unroll
  for (int i = 6; i >= 0; i--) v = v * x + c.c[i];
  return v;
}

__device__ double inv_norm_cdf_d(double p) {
  if (!(p > 0.0)) return -INFINITY;
  if (!(p < 1.0)) return INFINITY;
  const double q = p - 0.5;
  if (fabs(q) <= 0.425) {
    const double r = 0.180625 - q * q;
    return q * horner8(kAS241A, r) / horner8(kAS241B, r);
  }
  double r = sqrt(-log(q < 0.0 ? p : 1.0 - p));
  double x;
  if (r <= 5.0) {
    r -= 1.6;
    x = horner8(kAS241C, r) / horner8(kAS241D, r);
  } else {
    r -= 5.0;
    x = horner8(kAS241E, r) / horner8(kAS241F, r);
  }
  return q < 0.0 ? -x : x;
}

__device__ __forceinline__ double ud53(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)(a >> 6) << 27) | (uint64_t)(b >> 5);
  return ((double)k + 0.5) * (1.0 / 9007199254740992.0);
}

// Both branches start from block (c0, 0, it, tag), computed once before
// the branch so a wave whose lanes take both pays for one Philox.
__device__ double truncnorm0(double mean, uint64_t seed, uint32_t c0, uint32_t it, uint32_t tag) {
  const double abar = -mean;
  uint32_t w[4];
  philox_at(seed, c0, 0u, it, tag, w);
  if (abar > 3.0) {
    for (uint32_t t = 0; t < 64; t++) {
      if (t > 0) philox_at(seed, c0, t, it, tag, w);
      const double u = ud53(w[0], w[1]), v = ud53(w[2], w[3]);
      const double xbar = sqrt(abar * abar - 2.0 * log(1.0 - u));
      if (v <= xbar / abar) return xbar + mean;
    }
    return abar + mean;
  }
  const double lo = norm_cdf_d(abar);
  return inv_norm_cdf_d(lo + (1.0 - lo) * ud53(w[0], w[1])) + mean;
}

constexpr int kProbitThreads = 256;

// One wave handles 64 consecutive observations.  The factor rows are read
// by G = KP/4 lanes per observation (one float4 each, so every load
// instruction reads whole 128-byte lines), in G rounds of 64/G
// observations; each round's dot products are reduced across the lane
// group and handed to the lane that owns the observation, which then draws
// its z in FP64.
template <int KP>
__global__ void __launch_bounds__(kProbitThreads) k_probit_z(ProbitZArgs a) {
  constexpr int G = KP / 4;      // lanes per observation in the gather
  constexpr int PER = 64 / G;    // observations per round
  const uint32_t it = (uint32_t)*a.iter;
  const uint32_t tag = make_tag(TAG_PROBIT, (uint32_t)a.mode);
  const int lane = threadIdx.x & 63;
  const long long nchunk = (a.nnz + 63) / 64;
  const long long wave0 = ((long long)blockIdx.x * kProbitThreads + threadIdx.x) >> 6;
  const long long nwave = ((long long)gridDim.x * kProbitThreads) >> 6;
  for (long long chunk = wave0; chunk < nchunk; chunk += nwave) {
    const long long e = chunk * 64 + lane;
    const bool live = e < a.nnz;
    const int row = live ? a.row_of[e] : 0;
    const int col = live ? a.idx[e] : 0;
    float dot = 0.f;
#pragma unroll
    for (int r = 0; r < G; r++) {
      const int src = r * PER + lane / G;  // lane owning this round's observation
      const int rr = __shfl(row, src, 64), cc = __shfl(col, src, 64);
      const float4 x = reinterpret_cast<const float4*>(a.U + (size_t)rr * KP)[lane % G];
      const float4 w = reinterpret_cast<const float4*>(a.V + (size_t)cc * KP)[lane % G];
      float d = x.x * w.x + x.y * w.y + (x.z * w.z + x.w * w.w);
#pragma unroll
      for (int off = 1; off < G; off <<= 1) d += __shfl_xor(d, off, G);
      const float got = __shfl(d, (lane % PER) * G, 64);
      if (lane / PER == r) dot = got;
    }
    if (live) {
      const double y = 2.0 * a.y[e] - 1.0;
      a.z[e] = (float)(y * truncnorm0(y * (double)dot, a.seed, (uint32_t)e, it, tag));
    }
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
