// sampler_batched128.hip -- the batched per-row sampler for K in (64, 128]:
// split-BF16 MFMA Gram of a 128 x 128 precision matrix, one row per
// 256-thread work-group.
//
// Same mathematics as the K <= 64 samplers (sample_latent_blas,
// latentprior.cpp:361-385): P = Lambda + alpha G, b = alpha c + Lambda mu,
// u = L^{-T}(L^{-1} b + z); rows above the segment length are split into
// segments whose partial Grams are summed in segment order; light rows are
// taken by decreasing nnz.  What changes with K = 128:
//   * Gram: per 16-nonzero chunk, wave w gathers components 32w..32w+31
//     (lane l: component 32w + l%32 of the chunk's nonzeros 8(l/32)..+7,
//     the MFMA operand layout of v_mfma_f32_32x32x16_bf16), splits them into
//     BF16 hi + lo and stores them in LDS; the ten 32 x 32 blocks (I, J),
//     I >= J, of G are then shared out, each block taking the operands of
//     component blocks I and J from LDS (HiHi^T + HiLo^T + LoHi^T + LoLo^T):
//       wave 0: (0,0) (1,0) (1,1)    wave 1: (2,0) (2,1) (2,2)
//       wave 2: (3,0) (3,1)          wave 3: (3,2) (3,3)
//     The chunk's gathers are issued before the previous chunk's MFMAs, and
//     its index loads one chunk earlier still.
//   * Epilogue: each wave keeps its blocks of -P in the accumulator layout
//     and the matrix is factored in 4-column steps as the K = 64 epilogue
//     does within one wave: the owners of the step's block column publish
//     its 4 columns to LDS (double-buffered by step parity); lanes of waves
//     0 and 1 own rows 0-63 and 64-127 and each forms the 4 x 4 diagonal
//     factor, its own row of L, the forward solve and its multipliers, which
//     go to LDS; after a second barrier every wave applies the rank-4
//     update to its blocks that still hold unfactored columns on
//     v_mfma_f32_32x32x2_f32, its operands read from LDS (at most 6 MFMAs
//     per wave and step).  L is kept as a packed triangle in LDS (over the
//     Gram operands' buffers, three work-groups per CU); wave 0
//     does the back solve with two rows per lane (readlane broadcasts).
//   * No condition probes: refinement is not offered at this size.
#include <stdexcept>

#include "batched.h"

namespace mgpu {

constexpr int kPartial128 = 10 * 1024 + 256;  // floats per segment partial: 10 blocks + 4 x 64 RHS partials
constexpr int kThreads128 = 256;

// Gram blocks of each wave (block row, block column; -1: none)
__constant__ int kBI128[4][3] = {{0, 1, 1}, {2, 2, 2}, {3, 3, -1}, {3, 3, -1}};
__constant__ int kBJ128[4][3] = {{0, 0, 1}, {0, 1, 2}, {0, 1, -1}, {2, 3, -1}};

struct Lds128 {
  union {
    uint4 op[2][4][64][2];    // Gram operands: [buffer][component block][lane][hi, lo]
    float Lf[128 * 129 / 2];  // the epilogue's L, row i at i (i + 1) / 2 (the operands are dead by then)
  } u;
  float csh[128];           // RHS sums c (alpha c + Lambda mu is the RHS)
  float WT[2][4][128];      // the step's columns of -P (double-buffered by step parity)
  float Wb[2][128];         // the step's RHS values
  float qsh[4][128];        // the step's multipliers of every row (zero at and above the block)
  float hb[2][64];          // rows 64-127: b and 1 / L_ii, for the back solve on wave 0
  float us[128];            // the sampled row (SSE)
  float ws[64];             // rows 64-127: L^{-1} b + z (SSE)
  double sse[2];
};

struct Gram128 {
  f32x16 g[3];  // this wave's blocks
  float bp;     // RHS partial of component 32w + l%32 over this lane's half of each chunk
};

__device__ __forceinline__ void zero_gram(Gram128& g) {
#pragma unroll
  for (int b = 0; b < 3; b++)
#pragma unroll
    for (int r = 0; r < 16; r++) g.g[b][r] = 0.f;
  g.bp = 0.f;
}

// Gram stream over the nonzeros [beg, end) of one row (or segment); every
// wave of the work-group runs it with the same range.
template <int NO>
__device__ __forceinline__ void gram128(const SampleArgs& a, const Factors<NO>& F, Lds128& s, int beg, int end,
                                        Gram128& g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5;
  const int comp = 32 * w + (lane & 31);
  const int nch = (end - beg + 15) >> 4;
  const int I0 = kBI128[w][0], J0 = kBJ128[w][0], I1 = kBI128[w][1], J1 = kBJ128[w][1];
  const int I2 = kBI128[w][2], J2 = kBJ128[w][2];
  // indices and residual of nonzero e0 + lane, lanes 0..15 (index -1: none)
  auto load_ix = [&](int e0, int (&ix)[NO], float& r) {
    const int e = e0 + lane;
    const bool ok = lane < 16 && e < end;
#pragma unroll
    for (int t = 0; t < NO; t++) ix[t] = ok ? a.rows.idx[(size_t)e * NO + t] : -1;
    r = ok ? a.rows.r[e] : 0.f;
  };
  // this lane's 8 values of the chunk (component comp of nonzeros 8h..8h+7)
  auto gather = [&](const int (&ix)[NO], float (&v)[8]) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
      float x = ld_factor<128>(F.rs[0], __shfl(ix[0], 8 * h + t, 64), comp);
#pragma unroll
      for (int o = 1; o < NO; o++) x *= ld_factor<128>(F.rs[o], __shfl(ix[o], 8 * h + t, 64), comp);
      v[t] = x;
    }
  };
  if (nch == 0) return;
  int ixc[NO], ixn[NO];
  float rc, rn;
  float v[8];
  load_ix(beg, ixc, rc);
  gather(ixc, v);
  load_ix(beg + 16, ixn, rn);
  for (int c = 0; c < nch; c++) {
    const int buf = c & 1;
#pragma unroll
    for (int t = 0; t < 8; t++) g.bp += v[t] * __shfl(rc, 8 * h + t, 64);
    bf16x8 hi, lo;
    split8(v, hi, lo);
    s.u.op[buf][w][lane][0] = __builtin_bit_cast(uint4, hi);
    s.u.op[buf][w][lane][1] = __builtin_bit_cast(uint4, lo);
    __syncthreads();
    // the next chunk's gathers (and the one after's indices) before this
    // chunk's MFMAs
    float vn[8];
    if (c + 1 < nch) {
      gather(ixn, vn);
#pragma unroll
      for (int t = 0; t < NO; t++) ixc[t] = ixn[t];
      rc = rn;
      load_ix(beg + 16 * (c + 2), ixn, rn);
    }
    auto blk = [&](f32x16& acc, int I, int J) {
      const bf16x8 ha = __builtin_bit_cast(bf16x8, s.u.op[buf][I][lane][0]);
      const bf16x8 la = __builtin_bit_cast(bf16x8, s.u.op[buf][I][lane][1]);
      const bf16x8 hb = __builtin_bit_cast(bf16x8, s.u.op[buf][J][lane][0]);
      const bf16x8 lb = __builtin_bit_cast(bf16x8, s.u.op[buf][J][lane][1]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha, hb, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha, lb, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la, hb, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la, lb, acc, 0, 0, 0);
    };
    blk(g.g[0], I0, J0);
    blk(g.g[1], I1, J1);
    if (I2 >= 0) blk(g.g[2], I2, J2);
    if (c + 1 < nch) {
#pragma unroll
      for (int t = 0; t < 8; t++) v[t] = vn[t];
    }
  }
}

// The row's sample (and its share of the train SSE when sse_partial is set)
// from the work-group's Gram state; returns the SSE share on thread 0.
__device__ double finish128(const SampleArgs& a, int row, Gram128 g, int it, float alpha, Lds128& s) {
  int lane = threadIdx.x & 63;
  opaque(lane);
  const int w = threadIdx.x >> 6, half = lane >> 5, c32 = lane & 31;
  const bool active = row >= 0;
  // RHS sums c of the 128 components
  {
    const float c = g.bp + __shfl_xor(g.bp, 32, 64);
    if (lane < 32) s.csh[32 * w + lane] = c;
  }
  // each wave keeps its Gram blocks as -(Lambda + alpha G)
  const int nb = w < 2 ? 3 : 2;
  int BI[3], BJ[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    BI[q] = kBI128[w][q];
    BJ[q] = kBJ128[w][q];
  }
  f32x16 e[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    e[q] = g.g[q];
    if (q >= nb) continue;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const float4 l4 =
          *reinterpret_cast<const float4*>(a.Lambda_f + (32 * BJ[q] + c32) * 128 + 32 * BI[q] + 8 * m + 4 * half);
      const float el[4] = {l4.x, l4.y, l4.z, l4.w};
#pragma unroll
      for (int t = 0; t < 4; t++) e[q][4 * m + t] = -fmaf(alpha, e[q][4 * m + t], el[t]);
    }
  }
  __syncthreads();  // csh complete
  const int i = 64 * w + lane;  // this lane's row (waves 0 and 1)
  const bool rl = w < 2;
  float c = 0.f, b = 0.f, inv_diag = 1.f, zi = 0.f;
  bool bad = false;
  if (rl) {
    c = s.csh[i];
    b = alpha * c;
    if (a.Lmu_rows) {
      if (active) b += a.Lmu_rows[(size_t)row * 128 + i];
    } else {
      b += a.Lmu_f[i];
    }
    zi = i < a.K ? normal_f32(a.seed, (uint32_t)(active ? row : 0), (uint32_t)i, (uint32_t)it,
                              make_tag(TAG_LATENT, a.mode))
                 : 0.f;
  }
  const int tri_i = i * (i + 1) / 2;
#pragma unroll 1
  for (int k0 = 0; k0 < 128; k0 += 4) {
    opaque(lane);
    const int J0 = k0 >> 5, cc = k0 & 31, pb = (k0 >> 2) & 1;
    // 1. publish columns k0..k0+3 of -P (the owners of block column J0) and
    // the RHS
    if ((unsigned)(c32 - cc) < 4u) {
      const int u = c32 - cc;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        if (q >= nb || BJ[q] != J0) continue;
#pragma unroll
        for (int m = 0; m < 4; m++)
          *reinterpret_cast<float4*>(&s.WT[pb][u][32 * BI[q] + 8 * m + 4 * half]) =
              make_float4(e[q][4 * m], e[q][4 * m + 1], e[q][4 * m + 2], e[q][4 * m + 3]);
      }
    }
    if (rl) s.Wb[pb][i] = b;
    __syncthreads();
    // 2. the 4 x 4 diagonal factor, this row's entries of L and the forward
    // solve; the rows' multipliers go to LDS for the update
    if (rl) {
      const float w0 = -s.WT[pb][0][i], w1 = -s.WT[pb][1][i], w2 = -s.WT[pb][2][i], w3 = -s.WT[pb][3][i];
      const float4 c0 = *reinterpret_cast<const float4*>(&s.WT[pb][0][k0]);
      const float4 c1 = *reinterpret_cast<const float4*>(&s.WT[pb][1][k0]);
      const float4 c2 = *reinterpret_cast<const float4*>(&s.WT[pb][2][k0]);
      const float4 c3 = *reinterpret_cast<const float4*>(&s.WT[pb][3][k0]);
      const float4 bd = *reinterpret_cast<const float4*>(&s.Wb[pb][k0]);
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
        float* lr = s.u.Lf + tri_i + k0;
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
      if (k0 + 4 < 128) {
        s.qsh[0][i] = below ? p0 : 0.f;
        s.qsh[1][i] = below ? p1 : 0.f;
        s.qsh[2][i] = below ? p2 : 0.f;
        s.qsh[3][i] = below ? p3 : 0.f;
      }
    }
    if (k0 + 4 >= 128) break;
    __syncthreads();
    // 3. -P += p p^T over this wave's blocks that still hold unfactored
    // columns; the MFMA operand of block row X on lane l is the multiplier
    // l / 32 (or 2 + l / 32) of row 32 X + l % 32
    const int jmin = (k0 + 4) >> 5;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      if (q >= nb || BJ[q] < jmin) continue;
      const float a01 = s.qsh[half][32 * BI[q] + c32], a23 = s.qsh[2 + half][32 * BI[q] + c32];
      const float b01 = s.qsh[half][32 * BJ[q] + c32], b23 = s.qsh[2 + half][32 * BJ[q] + c32];
      e[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a01, b01, e[q], 0, 0, 0);
      e[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a23, b23, e[q], 0, 0, 0);
    }
  }
  if (bad && active && i == 0) atomicMax(a.fail, row + 1);  // the pivots are the same on every row lane
  b += zi;
  const float wv = b;  // L^{-1} b + z
  // back solve L^T u = w on wave 0, lane l holding rows l and 64 + l
  if (w == 1) {
    s.hb[0][lane] = b;
    s.hb[1][lane] = inv_diag;
    s.ws[lane] = wv;
  }
  __syncthreads();
  double sse = 0.0;
  if (w == 0) {
    float bA = b, iA = inv_diag, bB = s.hb[0][lane], iB = s.hb[1][lane];
    for (int gq = 127; gq >= 64; gq -= 8) {
      opaque(lane);
      float La[8], Lb[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int j = gq - u;
        La[u] = s.u.Lf[j * (j + 1) / 2 + lane];
        Lb[u] = s.u.Lf[j * (j + 1) / 2 + ((64 + lane) <= j ? 64 + lane : j)];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int j = gq - u;
        const float xj = rdlane(bB * iB, j - 64);
        bA -= La[u] * xj;
        bB = (64 + lane < j) ? bB - Lb[u] * xj : (64 + lane == j ? xj : bB);
      }
    }
    for (int gq = 63; gq >= 0; gq -= 8) {
      opaque(lane);
      float La[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int j = gq - u;
        La[u] = s.u.Lf[j * (j + 1) / 2 + (lane <= j ? lane : j)];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int j = gq - u;
        const float xj = rdlane(bA * iA, j);
        bA = (lane < j) ? bA - La[u] * xj : (lane == j ? xj : bA);
      }
    }
    if (active) {
      a.U[(size_t)row * 128 + lane] = bA;
      a.U[(size_t)row * 128 + 64 + lane] = bB;
    }
    if (a.sse_partial) {
      s.us[lane] = bA;
      s.us[64 + lane] = bB;
      lds_wait();
      // (w^2 - u Lambda u) / alpha - 2 u c per component, plus the row's
      // sum of squared residuals
      float t = 0.f;
#pragma unroll
      for (int hh = 0; hh < 2; hh++) {
        const int ci = 64 * hh + lane;
        const float* Lrow = a.Lambda_f + ci * 128;
        float lu = 0.f;
#pragma unroll 8
        for (int q = 0; q < 32; q++) {
          const float4 u4 = *reinterpret_cast<const float4*>(&s.us[4 * q]);
          const float4 l4 = *reinterpret_cast<const float4*>(Lrow + 4 * q);
          lu += l4.x * u4.x + l4.y * u4.y + l4.z * u4.z + l4.w * u4.w;
        }
        const float ww = hh ? s.ws[lane] : wv;
        const float uu = hh ? bB : bA;
        const float cc = s.csh[ci];
        t += (ww * ww - uu * lu) / alpha - 2.f * uu * cc;
      }
      if (active) {
        if (lane == 0) t += a.r2[row];
        sse = (double)t;
      }
    }
  }
  __syncthreads();  // the LDS of this row is dead for the next
  return sse;
}

// work-group SSE sum into sse_partial[slot] (wave 0 holds the shares)
__device__ inline void write_sse128(const SampleArgs& a, double acc, int slot) {
  if (!a.sse_partial) return;
  if (threadIdx.x < 64) {
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (threadIdx.x == 0) a.sse_partial[slot] = acc;
  }
}

template <int NO>
__global__ void __launch_bounds__(kThreads128, 3) k_light128(SampleArgs a, BatchPlan p) {
  __shared__ Lds128 s;
  const Factors<NO> F = make_factors<NO, 128>(a.others);
  const int it = *a.iter;
  const float alpha = (float)(*a.alpha);
  double sse = 0.0;
  for (int q = blockIdx.x; q < p.n_pair; q += gridDim.x) {
    const int4 rng = p.pair_rng[q];
    const int row = p.pair_row[q].x;
    Gram128 g;
    zero_gram(g);
    gram128<NO>(a, F, s, rng.x, rng.y, g);
    sse += finish128(a, row, g, it, alpha, s);
  }
  write_sse128(a, sse, blockIdx.x);
}

template <int NO>
__global__ void __launch_bounds__(kThreads128) k_seg128(SampleArgs a, BatchPlan p) {
  __shared__ Lds128 s;
  const Factors<NO> F = make_factors<NO, 128>(a.others);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int sg = blockIdx.x; sg < p.n_seg; sg += gridDim.x) {
    Gram128 g;
    zero_gram(g);
    gram128<NO>(a, F, s, p.seg_beg[sg], p.seg_end[sg], g);
    float* out = p.partial + (size_t)sg * kPartial128;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int I = kBI128[w][q], J = kBJ128[w][q];
      if (I < 0) continue;
      const int t = I * (I + 1) / 2 + J;
#pragma unroll
      for (int r = 0; r < 16; r++) out[t * 1024 + r * 64 + lane] = g.g[q][r];
    }
    out[10240 + w * 64 + lane] = g.bp;
    __syncthreads();  // the operand buffers are reused by the next segment
  }
}

__global__ void __launch_bounds__(kThreads128, 3) k_heavy128(SampleArgs a, BatchPlan p, int sse_slot0) {
  __shared__ Lds128 s;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int it = *a.iter;
  const float alpha = (float)(*a.alpha);
  double sse = 0.0;
  for (int q = blockIdx.x; q < p.n_heavy; q += gridDim.x) {
    // the slot's summed partial (k_seg_sum leaves it in the first segment)
    const float* in = p.partial + (size_t)p.heavy_seg0[q] * kPartial128;
    Gram128 g;
    zero_gram(g);
#pragma unroll
    for (int qq = 0; qq < 3; qq++) {
      const int I = kBI128[w][qq], J = kBJ128[w][qq];
      if (I < 0) continue;
      const int t = I * (I + 1) / 2 + J;
#pragma unroll
      for (int r = 0; r < 16; r++) g.g[qq][r] = in[t * 1024 + r * 64 + lane];
    }
    g.bp = in[10240 + w * 64 + lane];
    sse += finish128(a, p.heavy[q], g, it, alpha, s);
  }
  write_sse128(a, sse, sse_slot0 + blockIdx.x);
}

template <int NO>
static void launch128_t(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg) {
  const int gl = batched_light_grid(p), gh = batched_heavy_grid(p);
  const bool fork = p.n_seg > 0 && seg.stream;
  if (p.n_seg > 0) {
    const hipStream_t ss = fork ? seg.stream : s;
    if (fork) seg_fork(s, seg);
    hipLaunchKernelGGL((k_seg128<NO>), dim3(batched_seg_grid(p)), dim3(kThreads128), 0, ss, a, p);
    launch_seg_sum<kPartial128>(p, ss);
    if (gh > 0) hipLaunchKernelGGL(k_heavy128, dim3(gh), dim3(kThreads128), 0, ss, a, p, gl);
    if (fork) seg_record_join(seg);
  }
  if (gl > 0) hipLaunchKernelGGL((k_light128<NO>), dim3(gl), dim3(kThreads128), 0, s, a, p);
  if (fork) seg_wait_join(s, seg);
}

void launch_sample_batched128(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg) {
  switch (a.rows.nother) {
    case 1: return launch128_t<1>(a, p, s, seg);
    case 2: return launch128_t<2>(a, p, s, seg);
    case 3: return launch128_t<3>(a, p, s, seg);
    case 4: return launch128_t<4>(a, p, s, seg);
    default: break;
  }
}

}  // namespace mgpu
