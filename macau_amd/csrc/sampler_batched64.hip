// sampler_batched64.hip -- the batched per-row sampler for K in (32, 64]:
// split-BF16 MFMA Gram of a 64 x 64 precision matrix, one row per wave.
//
// Same mathematics and work decomposition as the K <= 32 sampler
// (sampler_batched.hip): per row P = Lambda + alpha G, b = alpha c + Lambda mu,
// u = L^{-T}(L^{-1} b + z); rows above 2048 nonzeros are split into segments
// whose partial Grams are summed in segment order; light rows are sorted by
// nnz.  What changes with K = 64:
//   * a lane gathers components l%32 and l%32 + 32 of its 8 nonzeros per
//     16-nonzero chunk; the Gram is three 32 x 32 accumulator blocks
//     G00 (rows/cols 0-31), G10 (rows 32-63 x cols 0-31), G11 (rows/cols
//     32-63), each HiHi^T + HiLo^T + LoHi^T + LoLo^T (12 v_mfma_f32_32x32x16_bf16
//     per chunk); the stream stages 32-nonzero blocks (index loads two blocks and
//     gathers one block ahead of the MFMAs);
//   * the epilogue factors one 64 x 64 matrix per wave (lane i = row i) in
//     4-column blocks with the trailing rank-4 updates on
//     v_mfma_f32_32x32x2_f32 over the three blocks (only G11 once the first
//     32 columns are done).
#include <stdexcept>

#include "batched.h"

namespace mgpu {

constexpr int kBlk64 = 32;                 // nonzeros per staged block (lanes 0..31 load)
constexpr int kPartial64 = 3 * 1024 + 128;  // floats per segment partial: 3 fragments + 2 x 64 RHS

template <int NO>
struct BlkLds64 {
  int ix[2][NO][kBlk64];  // double buffered
  float r[2][kBlk64];
};

struct Gram64 {
  f32x16 g00, g10, g11;  // blocks of G in the accumulator layout
  float b0, b1;          // RHS partials of components l%32 and l%32 + 32
};

__device__ __forceinline__ void zero_gram(Gram64& g) {
#pragma unroll
  for (int r = 0; r < 16; r++) {
    g.g00[r] = 0.f;
    g.g10[r] = 0.f;
    g.g11[r] = 0.f;
  }
  g.b0 = 0.f;
  g.b1 = 0.f;
}

template <int NO>
__device__ __forceinline__ void load_blk64(const SampleArgs& a, int e0, int end, int lane, BlkReg<NO>& b) {
  const int e = e0 + lane;
  if (lane < kBlk64 && e < end) {
#pragma unroll
    for (int t = 0; t < NO; t++) b.ix[t] = a.rows.idx[(size_t)e * NO + t];
    b.r = a.rows.r[e];
  } else {
#pragma unroll
    for (int t = 0; t < NO; t++) b.ix[t] = -1;
    b.r = 0.f;
  }
}

template <int NO>
__device__ __forceinline__ void stage_blk64(BlkLds64<NO>& L, int buf, int lane, const BlkReg<NO>& b) {
  if (lane < kBlk64) {
#pragma unroll
    for (int t = 0; t < NO; t++) L.ix[buf][t][lane] = b.ix[t];
    L.r[buf][lane] = b.r;
  }
}

// components comp and comp + 32 of the 8 nonzeros of this lane's half in
// each 16-nonzero chunk of the block
template <int NO>
__device__ __forceinline__ void gather_blk64(const Factors<NO>& F, const BlkLds64<NO>& L, int buf, int half,
                                             int comp, float v0[16], float v1[16]) {
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int o = c * 16 + half * 8;
    int ix[NO][8];
#pragma unroll
    for (int t = 0; t < NO; t++)
#pragma unroll
      for (int i = 0; i < 8; i++) ix[t][i] = L.ix[buf][t][o + i];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      float x0 = ld_factor<64>(F.rs[0], ix[0][i], comp);
      float x1 = ld_factor<64>(F.rs[0], ix[0][i], comp + 32);
#pragma unroll
      for (int t = 1; t < NO; t++) {
        x0 *= ld_factor<64>(F.rs[t], ix[t][i], comp);
        x1 *= ld_factor<64>(F.rs[t], ix[t][i], comp + 32);
      }
      v0[c * 8 + i] = x0;
      v1[c * 8 + i] = x1;
    }
  }
}

template <int NO>
__device__ __forceinline__ void mfma_blk64(const BlkLds64<NO>& L, int buf, int half, const float v0[16],
                                           const float v1[16], int nvalid, Gram64& g) {
#pragma unroll
  for (int c = 0; c < 2; c++) {
    if (c * 16 >= nvalid) break;
    const int o = c * 16 + half * 8;
    bf16x8 h0, l0, h1, l1;
    split8(v0 + c * 8, h0, l0);
    split8(v1 + c * 8, h1, l1);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float r = L.r[buf][o + i];
      g.b0 += v0[c * 8 + i] * r;
      g.b1 += v1[c * 8 + i] * r;
    }
    g.g00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h0, h0, g.g00, 0, 0, 0);
    g.g00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h0, l0, g.g00, 0, 0, 0);
    g.g00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(l0, h0, g.g00, 0, 0, 0);
    g.g00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(l0, l0, g.g00, 0, 0, 0);
    g.g10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h1, h0, g.g10, 0, 0, 0);
    g.g10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h1, l0, g.g10, 0, 0, 0);
    g.g10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(l1, h0, g.g10, 0, 0, 0);
    g.g10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(l1, l0, g.g10, 0, 0, 0);
    g.g11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h1, h1, g.g11, 0, 0, 0);
    g.g11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h1, l1, g.g11, 0, 0, 0);
    g.g11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(l1, h1, g.g11, 0, 0, 0);
    g.g11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(l1, l1, g.g11, 0, 0, 0);
  }
}

// Gram and RHS partials of nonzeros [beg, end) of one row (or segment)
template <int NO>
struct GramStream64 {
  const SampleArgs& a;
  const Factors<NO>& F;
  BlkLds64<NO>& L;
  int beg, end, nb, lane, half, comp;
  BlkReg<NO> nxt;

  __device__ __forceinline__ GramStream64(const SampleArgs& a_, const Factors<NO>& F_, BlkLds64<NO>& L_, int b,
                                          int e)
      : a(a_), F(F_), L(L_), beg(b), end(e) {
    lane = threadIdx.x & 63;
    half = lane >> 5;
    comp = lane & 31;
    nb = (end - beg + kBlk64 - 1) / kBlk64;
  }

  __device__ __forceinline__ void step(int k, const float c0[16], const float c1[16], float n0[16], float n1[16],
                                       Gram64& g) {
    BlkReg<NO> after;
    if (k + 2 < nb) load_blk64<NO>(a, beg + (k + 2) * kBlk64, end, lane, after);
    if (k + 1 < nb) {
      stage_blk64<NO>(L, (k + 1) & 1, lane, nxt);
      lds_wait();
      gather_blk64<NO>(F, L, (k + 1) & 1, half, comp, n0, n1);
    }
    mfma_blk64<NO>(L, k & 1, half, c0, c1, end - (beg + k * kBlk64), g);
    nxt = after;
  }

  // index/residual loads of the first two blocks of [b, e), issued ahead
  // of run() by the caller
  static __device__ __forceinline__ void first_blocks(const SampleArgs& a, int b, int e, int lane,
                                                      BlkReg<NO>& f0, BlkReg<NO>& f1) {
    load_blk64<NO>(a, b, e, lane, f0);
    load_blk64<NO>(a, b + kBlk64, e, lane, f1);
  }

  __device__ __forceinline__ void run(const BlkReg<NO>& f0, const BlkReg<NO>& f1, Gram64& g) {
    if (nb == 0) return;
    float a0[16], a1[16], b0[16], b1[16];
    nxt = f1;
    stage_blk64<NO>(L, 0, lane, f0);
    lds_wait();
    gather_blk64<NO>(F, L, 0, half, comp, a0, a1);
    for (int k = 0; k < nb; k += 2) {
      step(k, a0, a1, b0, b1, g);
      if (k + 1 < nb) step(k + 1, b0, b1, a0, a1, g);
    }
  }
};

struct Fin64Lds {
  alignas(16) float WT[4][64];     // current block columns of -P (column-major)
  alignas(16) float Wb[4][64];     // RHS and the three probe vectors
  alignas(16) float Lf[64 * 65 / 2];  // L, lower triangle, row i at i (i + 1) / 2
  alignas(16) float us[64];        // the new sample (SSE)
};

// Finish one row whose Gram is in g (row < 0: no row); returns this lane's
// SSE share.  Lane i owns row i of P.  With refinement on, three probe
// vectors (z, all ones, alternating signs) are forward-solved beside the RHS
// for the condition estimate (probe_estimate).
__device__ float finish64(const SampleArgs& a, int row, Gram64 g, int it, float alpha, Fin64Lds& s, int nnz) {
  int lane = threadIdx.x & 63;
  opaque(lane);
  const int half = lane >> 5, c32 = lane & 31;
  int i = lane;
  const bool active = row >= 0;
  // -P = -(Lambda + alpha G) in the accumulator layout: block (bi, bj), lane l,
  // register 4m + t is entry (32 bi + 8m + 4h + t, 32 bj + l%32)
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const float4 l00 = *reinterpret_cast<const float4*>(a.Lambda_f + c32 * 64 + 8 * m + 4 * half);
    const float4 l10 = *reinterpret_cast<const float4*>(a.Lambda_f + c32 * 64 + 32 + 8 * m + 4 * half);
    const float4 l11 = *reinterpret_cast<const float4*>(a.Lambda_f + (32 + c32) * 64 + 32 + 8 * m + 4 * half);
    const float e00[4] = {l00.x, l00.y, l00.z, l00.w};
    const float e10[4] = {l10.x, l10.y, l10.z, l10.w};
    const float e11[4] = {l11.x, l11.y, l11.z, l11.w};
#pragma unroll
    for (int t = 0; t < 4; t++) {
      g.g00[4 * m + t] = -fmaf(alpha, g.g00[4 * m + t], e00[t]);
      g.g10[4 * m + t] = -fmaf(alpha, g.g10[4 * m + t], e10[t]);
      g.g11[4 * m + t] = -fmaf(alpha, g.g11[4 * m + t], e11[t]);
    }
  }
  // RHS of component i: partials of both halves of the wave
  const float s0 = __shfl_xor(g.b0, 32, 64), s1 = __shfl_xor(g.b1, 32, 64);
  const float c = half ? (g.b1 + s1) : (g.b0 + s0);
  float b = alpha * c;
  if (a.Lmu_rows) {
    if (active) b += a.Lmu_rows[(size_t)row * 64 + i];
  } else {
    b += a.Lmu_f[i];
  }
  bool bad = false;
  float inv_diag = 1.f;
  const float zi = i < a.K ? normal_f32(a.seed, (uint32_t)(active ? row : 0), (uint32_t)i, (uint32_t)it,
                                        make_tag(TAG_LATENT, a.mode))
                           : 0.f;
  // condition probes (refinement only): L^{-1} D^{1/2} x for x = z, ones,
  // alternating signs, D = diag(P)
  const bool probe = a.refine.count != nullptr && nnz >= a.refine.min_nnz;
  float x1 = zi, x2 = i < a.K ? 1.f : 0.f, x3 = i < a.K ? ((i & 1) ? -1.f : 1.f) : 0.f;
  if (probe) {
    // row i < 32 of P is in block (0, 0), row i >= 32 in block (1, 1)
    const float d00 = frag_diag(g.g00, lane), d11 = frag_diag(g.g11, lane);
    const float sd = sqrtf(-(half ? d11 : d00));  // the fragments hold -P
    x1 *= sd;
    x2 *= sd;
    x3 *= sd;
  }
  const int tri_i = i * (i + 1) / 2;
#pragma unroll
  for (int k0 = 0; k0 < 64; k0 += 4) {
    opaque(i);
    // 1. publish columns k0..k0+3 of -P, rows from the block row of k0 down
    const int cc = k0 & 31;
    if ((unsigned)(c32 - cc) < 4u) {
      const int u = c32 - cc;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        if (k0 < 32) {
          if (8 * m + 8 > (k0 & ~7))
            *reinterpret_cast<float4*>(&s.WT[u][8 * m + 4 * half]) =
                make_float4(g.g00[4 * m], g.g00[4 * m + 1], g.g00[4 * m + 2], g.g00[4 * m + 3]);
          *reinterpret_cast<float4*>(&s.WT[u][32 + 8 * m + 4 * half]) =
              make_float4(g.g10[4 * m], g.g10[4 * m + 1], g.g10[4 * m + 2], g.g10[4 * m + 3]);
        } else if (32 + 8 * m + 8 > (k0 & ~7)) {
          *reinterpret_cast<float4*>(&s.WT[u][32 + 8 * m + 4 * half]) =
              make_float4(g.g11[4 * m], g.g11[4 * m + 1], g.g11[4 * m + 2], g.g11[4 * m + 3]);
        }
      }
    }
    s.Wb[0][i] = b;
    if (probe) {
      s.Wb[1][i] = x1;
      s.Wb[2][i] = x2;
      s.Wb[3][i] = x3;
    }
    lds_wait();
    const float w0 = -s.WT[0][i], w1 = -s.WT[1][i], w2 = -s.WT[2][i], w3 = -s.WT[3][i];
    const float4 c0 = *reinterpret_cast<const float4*>(&s.WT[0][k0]);
    const float4 c1 = *reinterpret_cast<const float4*>(&s.WT[1][k0]);
    const float4 c2 = *reinterpret_cast<const float4*>(&s.WT[2][k0]);
    const float4 c3 = *reinterpret_cast<const float4*>(&s.WT[3][k0]);
    const float4 bd = *reinterpret_cast<const float4*>(&s.Wb[0][k0]);
    const float a00 = -c0.x;
    bool bp = !(a00 > 0.f);
    const float r0 = __builtin_amdgcn_rsqf(bp ? 1.f : a00);
    const float l10 = -c0.y * r0, l20 = -c0.z * r0, l30 = -c0.w * r0;
    const float a11 = -c1.y - l10 * l10;
    bp = bp || !(a11 > 0.f);
    const float r1 = __builtin_amdgcn_rsqf(a11 > 0.f ? a11 : 1.f);
    const float l21 = (-c1.z - l20 * l10) * r1, l31 = (-c1.w - l30 * l10) * r1;
    const float a22 = -c2.z - l20 * l20 - l21 * l21;
    bp = bp || !(a22 > 0.f);
    const float r2 = __builtin_amdgcn_rsqf(a22 > 0.f ? a22 : 1.f);
    const float l32 = (-c2.w - l30 * l20 - l31 * l21) * r2;
    const float a33 = -c3.w - l30 * l30 - l31 * l31 - l32 * l32;
    bp = bp || !(a33 > 0.f);
    const float r3 = __builtin_amdgcn_rsqf(a33 > 0.f ? a33 : 1.f);
    bad = bad || bp;
    const float y0 = bd.x * r0;
    const float y1 = (bd.y - l10 * y0) * r1;
    const float y2 = (bd.z - l20 * y0 - l21 * y1) * r2;
    const float y3 = (bd.w - l30 * y0 - l31 * y1 - l32 * y2) * r3;
    const float p0 = w0 * r0;
    const float p1 = (w1 - p0 * l10) * r1;
    const float p2 = (w2 - p0 * l20 - p1 * l21) * r2;
    const float p3 = (w3 - p0 * l30 - p1 * l31 - p2 * l32) * r3;
    if (k0 <= i) {
      float* lr = s.Lf + tri_i + k0;
      lr[0] = p0;
      if (k0 + 1 <= i) lr[1] = p1;
      if (k0 + 2 <= i) lr[2] = p2;
      if (k0 + 3 <= i) lr[3] = p3;
    }
    const int t = i - k0;
    const bool inb = t >= 0 && t < 4, below = t >= 4;
    const float yt = t == 0 ? y0 : (t == 1 ? y1 : (t == 2 ? y2 : y3));
    const float rt = t == 0 ? r0 : (t == 1 ? r1 : (t == 2 ? r2 : r3));
    b = inb ? yt : (below ? b - (p0 * y0 + p1 * y1 + p2 * y2 + p3 * y3) : b);
    inv_diag = inb ? rt : inv_diag;
    if (probe) {
      // the same elimination step for the probe vectors
      auto fwd = [&](float x, float4 xd) {
        const float v0 = xd.x * r0;
        const float v1 = (xd.y - l10 * v0) * r1;
        const float v2 = (xd.z - l20 * v0 - l21 * v1) * r2;
        const float v3 = (xd.w - l30 * v0 - l31 * v1 - l32 * v2) * r3;
        const float vt = t == 0 ? v0 : (t == 1 ? v1 : (t == 2 ? v2 : v3));
        return inb ? vt : (below ? x - (p0 * v0 + p1 * v1 + p2 * v2 + p3 * v3) : x);
      };
      x1 = fwd(x1, *reinterpret_cast<const float4*>(&s.Wb[1][k0]));
      x2 = fwd(x2, *reinterpret_cast<const float4*>(&s.Wb[2][k0]));
      x3 = fwd(x3, *reinterpret_cast<const float4*>(&s.Wb[3][k0]));
    }
    // 2. -P += p p^T over the rows below the block
    if (k0 + 4 < 64) {
      const float q0 = below ? p0 : 0.f, q1 = below ? p1 : 0.f, q2 = below ? p2 : 0.f, q3 = below ? p3 : 0.f;
      // swap(old, src): .x = (q0 | q1 of rows 0-31), .y = (q0 | q1 of rows 32-63)
      const auto s01 = __builtin_amdgcn_permlane32_swap(__float_as_uint(q0), __float_as_uint(q1), false, false);
      const auto s23 = __builtin_amdgcn_permlane32_swap(__float_as_uint(q2), __float_as_uint(q3), false, false);
      const float t01 = __uint_as_float(s01[0]), u01 = __uint_as_float(s01[1]);
      const float t23 = __uint_as_float(s23[0]), u23 = __uint_as_float(s23[1]);
      if (k0 + 4 < 32) {
        g.g00 = __builtin_amdgcn_mfma_f32_32x32x2f32(t01, t01, g.g00, 0, 0, 0);
        g.g00 = __builtin_amdgcn_mfma_f32_32x32x2f32(t23, t23, g.g00, 0, 0, 0);
        g.g10 = __builtin_amdgcn_mfma_f32_32x32x2f32(u01, t01, g.g10, 0, 0, 0);
        g.g10 = __builtin_amdgcn_mfma_f32_32x32x2f32(u23, t23, g.g10, 0, 0, 0);
      }
      g.g11 = __builtin_amdgcn_mfma_f32_32x32x2f32(u01, u01, g.g11, 0, 0, 0);
      g.g11 = __builtin_amdgcn_mfma_f32_32x32x2f32(u23, u23, g.g11, 0, 0, 0);
    }
  }
  if (bad && active && i == 0) atomicMax(a.fail, row + 1);
  b += zi;
  const float w = b;  // L^{-1} b + z
  // back solve L^T u = w: column i of L read from the rows in LDS, eight
  // steps' worth at a time (lanes i > j read later rows, unused); readlane
  // broadcast of the solved values
  lds_wait();
#pragma unroll
  for (int g = 63; g >= 0; g -= 8) {
    opaque(i);
    float Lc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) Lc[u] = s.Lf[(g - u) * (g - u + 1) / 2 + i];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int j = g - u;
      const float xj = rdlane(b * inv_diag, j);
      b = (i < j) ? b - Lc[u] * xj : (i == j ? xj : b);
    }
  }
  if (active) a.U[(size_t)row * 64 + i] = b;
  if (probe) {
    const float est = probe_estimate<64>(a.K, x1, x2, x3, zi);
    refine_enqueue<64, 64>(a, active ? row : -1, nnz, i, 0, est, s.Lf + tri_i);
  }
  if (!a.sse_partial) return 0.f;
  s.us[i] = b;
  lds_wait();
  const float* Lrow = a.Lambda_f + i * 64;
  float lu = 0.f;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const float4 u4 = *reinterpret_cast<const float4*>(&s.us[4 * q]);
    const float4 l4 = *reinterpret_cast<const float4*>(Lrow + 4 * q);
    lu += l4.x * u4.x + l4.y * u4.y + l4.z * u4.z + l4.w * u4.w;
  }
  float t = 0.f;
  if (active) {
    t = (w * w - b * lu) / alpha - 2.f * b * c;
    if (i == 0) t += a.r2[row];
  }
  return t;
}

template <int NO>
__global__ void __launch_bounds__(64, 3) k_light64(SampleArgs a, BatchPlan p) {
  __shared__ Fin64Lds s;
  __shared__ BlkLds64<NO> blk;
  const Factors<NO> F = make_factors<NO, 64>(a.others);
  const int it = *a.iter;
  const float alpha = (float)(*a.alpha);
  const int lane = threadIdx.x;
  double sse = 0.0;
  // rows q, q + G, ...: the next row's nnz range and first two index blocks
  // are loaded while the current row is multiplied and finished
  int q = blockIdx.x;
  int4 rng = make_int4(0, 0, 0, 0);
  int row = -1;
  BlkReg<NO> f0, f1;
  if (q < p.n_pair) {
    rng = p.pair_rng[q];
    row = p.pair_row[q].x;
    GramStream64<NO>::first_blocks(a, rng.x, rng.y, lane, f0, f1);
  }
  for (; q < p.n_pair; q += gridDim.x) {
    const int qn = q + gridDim.x;
    int4 rng_n = make_int4(0, 0, 0, 0);
    int row_n = -1;
    if (qn < p.n_pair) {
      rng_n = p.pair_rng[qn];
      row_n = p.pair_row[qn].x;
    }
    Gram64 g;
    zero_gram(g);
    GramStream64<NO>(a, F, blk, rng.x, rng.y).run(f0, f1, g);
    if (qn < p.n_pair) GramStream64<NO>::first_blocks(a, rng_n.x, rng_n.y, lane, f0, f1);
    sse += (double)finish64(a, row, g, it, alpha, s, rng.y - rng.x);
    rng = rng_n;
    row = row_n;
  }
  write_sse(a, sse, blockIdx.x);
}

template <int NO>
__global__ void __launch_bounds__(64, 2) k_seg64(SampleArgs a, BatchPlan p) {
  __shared__ BlkLds64<NO> blk;
  const Factors<NO> F = make_factors<NO, 64>(a.others);
  const int lane = threadIdx.x;
  for (int sg = blockIdx.x; sg < p.n_seg; sg += gridDim.x) {
    Gram64 g;
    zero_gram(g);
    BlkReg<NO> f0, f1;
    GramStream64<NO>::first_blocks(a, p.seg_beg[sg], p.seg_end[sg], lane, f0, f1);
    GramStream64<NO>(a, F, blk, p.seg_beg[sg], p.seg_end[sg]).run(f0, f1, g);
    float* out = p.partial + (size_t)sg * kPartial64;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      out[r * 64 + lane] = g.g00[r];
      out[1024 + r * 64 + lane] = g.g10[r];
      out[2048 + r * 64 + lane] = g.g11[r];
    }
    out[3072 + lane] = g.b0;
    out[3136 + lane] = g.b1;
  }
}

__global__ void __launch_bounds__(64, 2) k_heavy64(SampleArgs a, BatchPlan p, int sse_slot0) {
  __shared__ Fin64Lds s;
  const int lane = threadIdx.x;
  const int it = *a.iter;
  const float alpha = (float)(*a.alpha);
  double sse = 0.0;
  for (int q = blockIdx.x; q < p.n_heavy; q += gridDim.x) {
    // the slot's summed partial (k_seg_sum leaves it in the first segment)
    Gram64 g;
    const float* in = p.partial + (size_t)p.heavy_seg0[q] * kPartial64;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      g.g00[r] = in[r * 64 + lane];
      g.g10[r] = in[1024 + r * 64 + lane];
      g.g11[r] = in[2048 + r * 64 + lane];
    }
    g.b0 = in[3072 + lane];
    g.b1 = in[3136 + lane];
    sse += (double)finish64(a, p.heavy[q], g, it, alpha, s, p.seg_len);
  }
  write_sse(a, sse, sse_slot0 + blockIdx.x);
}

int batched64_light_grid(const BatchPlan& p) {
  const int cap = p.light_cap > 0 ? p.light_cap : kDefaultLightCap;
  return p.n_pair < cap ? p.n_pair : cap;
}
int batched64_heavy_grid(const BatchPlan& p) {
  const int cap = p.heavy_cap > 0 ? p.heavy_cap : kDefaultHeavyCap;
  return p.n_heavy < cap ? p.n_heavy : cap;
}

template <int NO>
static void launch64_t(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg) {
  const int gl = batched64_light_grid(p), gh = batched64_heavy_grid(p);
  const bool fork = p.n_seg > 0 && seg.stream;
  if (p.n_seg > 0) {
    const hipStream_t ss = fork ? seg.stream : s;
    if (fork) seg_fork(s, seg);
    const int gs = batched_seg_grid(p);
    hipLaunchKernelGGL((k_seg64<NO>), dim3(gs), dim3(64), 0, ss, a, p);
    launch_seg_sum<kPartial64>(p, ss);
    // the heavy rows depend only on their segments: finished on the segment
    // stream too, beside the light rows
    if (gh > 0) hipLaunchKernelGGL(k_heavy64, dim3(gh), dim3(64), 0, ss, a, p, gl);
    if (fork) seg_record_join(seg);
  }
  if (gl > 0) hipLaunchKernelGGL((k_light64<NO>), dim3(gl), dim3(64), 0, s, a, p);
  if (fork) seg_wait_join(s, seg);
  launch_refine(a, s);
}

void launch_sample_batched64(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg) {
  switch (a.rows.nother) {
    case 1: return launch64_t<1>(a, p, s, seg);
    case 2: return launch64_t<2>(a, p, s, seg);
    case 3: return launch64_t<3>(a, p, s, seg);
    case 4: return launch64_t<4>(a, p, s, seg);
    default: break;
  }
}

}  // namespace mgpu
