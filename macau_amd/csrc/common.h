// common.h -- shared device/host definitions for the MI355X Gibbs sampler.
//
// Counter-based randomness: every draw is a pure function of
// (seed, c0, c1, iteration, tag), so a draw does not depend on which wave
// or thread computes it.  This replaces the reference's per-thread
// mt19937 streams (mvnormal.cpp:19-166), which are seeded from the clock
// and random_device and therefore unreproducible.
//
//   key      = (seed & 0xffffffff, seed >> 32)
//   normal j of stream (c0, it, tag): block philox(c0, j/4, it, tag),
//            Box-Muller on word pairs (w0,w1) -> z0,z1 and (w2,w3) -> z2,z3,
//            component j%4
//   gamma of stream (c0, it, tag): Marsaglia-Tsang, attempt t uses block
//            philox(c0, t, it, tag): normal from (w0,w1), uniform from w2
//   tag      = (stream << 24) | mode
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MGPU_HD __host__ __device__ __forceinline__

namespace mgpu {

enum : uint32_t {
  TAG_LATENT = 1u,       // c0 = row,  j = latent component
  TAG_WISHART_Z = 2u,    // c0 = i,    j = column of the Bartlett factor
  TAG_WISHART_CHI = 3u,  // c0 = i,    gamma for the i-th diagonal
  TAG_NW_MU = 4u,        // c0 = 0,    j = component of mu
  TAG_NOISE = 5u,        // c0 = 0,    gamma for the noise precision
  TAG_BETA_U = 6u,       // c0 = col,  j = component (link-matrix RHS)
  TAG_BETA_F = 7u,       // c0 = feat, j = component (link-matrix prior)
  TAG_LAMBDA_BETA = 8u,  // c0 = 0,    gamma for lambda_beta
  TAG_PROBIT = 9u,       // c0 = observation (mode order), c1 = attempt
  TAG_BETA_ONE = 10u,    // c0 = feature, j = latent (univariate link matrix)
};
MGPU_HD uint32_t make_tag(uint32_t stream, uint32_t mode) { return (stream << 24) | mode; }

// Philox4x32-10 (Salmon et al., SC'11).
MGPU_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                           uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
#else
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
#endif
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

MGPU_HD void philox_at(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t it, uint32_t tag, uint32_t w[4]) {
  philox4x32_10(c0, c1, it, tag, (uint32_t)seed, (uint32_t)(seed >> 32), w);
}

// 24-bit uniforms: exact in FP32 and FP64.
MGPU_HD float uf_open01(uint32_t x) { return ((float)(x >> 8) + 1.0f) * (1.0f / 16777216.0f); }
MGPU_HD float uf_co01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
MGPU_HD double ud_open01(uint32_t x) { return ((double)(x >> 8) + 1.0) * (1.0 / 16777216.0); }
MGPU_HD double ud_co01(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// FP32 Box-Muller of one word pair; sin/cos of 2*pi*u via sincospi (exact
// argument reduction).
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& z0, float& z1) {
  const float r = sqrtf(-2.0f * logf(uf_open01(a)));
  float s, c;
  sincospif(2.0f * uf_co01(b), &s, &c);
  z0 = r * c;
  z1 = r * s;
}
__device__ __forceinline__ void box_muller_f64(uint32_t a, uint32_t b, double& z0, double& z1) {
  const double r = sqrt(-2.0 * log(ud_open01(a)));
  double s, c;
  sincospi(2.0 * ud_co01(b), &s, &c);
  z0 = r * c;
  z1 = r * s;
}
// normal j of stream (c0, it, tag), FP32
__device__ __forceinline__ float normal_f32(uint64_t seed, uint32_t c0, uint32_t j, uint32_t it, uint32_t tag) {
  uint32_t w[4];
  philox_at(seed, c0, j >> 2, it, tag, w);
  float z0, z1;
  const int p = (j >> 1) & 1;
  box_muller_f32(w[2 * p], w[2 * p + 1], z0, z1);
  return (j & 1) ? z1 : z0;
}
__device__ __forceinline__ double normal_f64(uint64_t seed, uint32_t c0, uint32_t j, uint32_t it, uint32_t tag) {
  uint32_t w[4];
  philox_at(seed, c0, j >> 2, it, tag, w);
  double z0, z1;
  const int p = (j >> 1) & 1;
  box_muller_f64(w[2 * p], w[2 * p + 1], z0, z1);
  return (j & 1) ? z1 : z0;
}
// normals 4b .. 4b+3 of stream (c0, it, tag): the four values normal_f64
// returns for j = 4b .. 4b+3, from one Philox block
__device__ __forceinline__ void normal_f64x4(uint64_t seed, uint32_t c0, uint32_t b, uint32_t it, uint32_t tag,
                                             double z[4]) {
  uint32_t w[4];
  philox_at(seed, c0, b, it, tag, w);
  box_muller_f64(w[0], w[1], z[0], z[1]);
  box_muller_f64(w[2], w[3], z[2], z[3]);
}
// Gamma(shape, 1): Marsaglia & Tsang, ACM TOMS 26(3), 2000; shape < 1 by
// the boost Gamma(shape+1) * U^(1/shape), U from block (c0, 0xffffffff).
__device__ inline double gamma_f64(uint64_t seed, double shape, uint32_t c0, uint32_t it, uint32_t tag) {
  const double a = shape < 1.0 ? shape + 1.0 : shape;
  const double d = a - 1.0 / 3.0;
  const double c = 1.0 / sqrt(9.0 * d);
  double g = d;
  for (uint32_t t = 0; t < 32; t++) {
    uint32_t w[4];
    philox_at(seed, c0, t, it, tag, w);
    double x, x1;
    box_muller_f64(w[0], w[1], x, x1);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    const double u = ud_open01(w[2]);
    const double x2 = x * x;
    if (u < 1.0 - 0.0331 * x2 * x2 || log(u) < 0.5 * x2 + d * (1.0 - v + log(v))) {
      g = d * v;
      break;
    }
  }
  if (shape < 1.0) {
    uint32_t w[4];
    philox_at(seed, c0, 0xffffffffu, it, tag, w);
    g *= pow(ud_open01(w[0]), 1.0 / shape);
  }
  return g;
}

// ---- wave-level helpers (work-groups of one wave) ----

// LDS writes of this wave must land before its lanes read each other's
// entries: wait for the LDS queue (lgkmcnt(0), vmcnt/expcnt untouched).
__device__ __forceinline__ void lds_wait() {
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ double rdlane(double v, int l) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// An empty asm that "changes" v: predicates built from v after it are formed
// at that point instead of being hoisted and held in scalar registers.
__device__ __forceinline__ void opaque(int& v) { asm volatile("" : "+v"(v)); }
// Same, and v "depends" on dep: memory accesses addressed through v wait for
// dep to be computed (keeps a long unrolled chain from hoisting all its loads).
__device__ __forceinline__ void opaque_after(int& v, double dep) { asm volatile("" : "+v"(v) : "v"(dep)); }

// Padded latent dimension used for storage and kernel templates.
inline int pad_latent(int K) {
  if (K <= 8) return 8;
  if (K <= 16) return 16;
  if (K <= 32) return 32;
  if (K <= 64) return 64;
  return 128;
}
// The batched sampler's kernels tile 32 or 64 components: smaller latent
// sizes are stored padded to 32 (the padding components are zero).
inline int pad_latent_batched(int K) { return K <= 32 ? 32 : pad_latent(K); }

}  // namespace mgpu
