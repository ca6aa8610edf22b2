// sampler_batched.hip -- nnz-bucketed per-row sampler with a split-BF16
// MFMA Gram (K in (16, 32]).
//
// Same mathematics as sample_latent_blas (latentprior.cpp:361-385) and
// sample_latent_tensor (latentprior.cpp:118-161): per row
//   P = Lambda + alpha G,  G = sum_e c_e c_e^T,  b = alpha sum_e c_e r_e + Lambda mu_n,
//   u = L^{-T}(L^{-1} b + z),  P = L L^T.
//
// Work decomposition of one half-sweep (BatchPlan, built on the host):
//   * rows with more than 2048 nonzeros are split into 2048-nnz segments;
//     k_seg32 computes each segment's partial Gram and RHS (one wave per
//     segment), k_seg_sum adds each row's partials in segment order (one
//     thread per partial entry), and k_heavy32 finishes two rows per wave;
//   * the other rows, sorted by nnz (descending), are taken two per wave by
//     k_light32 (similar nnz in a pair keeps both halves busy).
// Gram: each MFMA step takes 16 nonzeros (lanes 0-31 the first 8, lanes
// 32-63 the next 8; lane l holds component l%32), splits every value into
// hi = bf16(x) and lo = bf16(x - hi), and accumulates G += Hi Hi^T + Hi Lo^T
// + Lo Hi^T + Lo Lo^T with four v_mfma_f32_32x32x16_bf16 into FP32.  Without
// Lo Lo^T the diagonal loses a positive bias of ~5e-6 |x|^2 per product,
// which rows of few nonzeros against a weak prior direction turn into row
// errors near 1e-4; the fourth MFMA costs no measurable time (the sampler is
// not MFMA-bound).  The RHS accumulates in FP32 on the vector ALU.
// The nonzeros stream through in blocks (32 nonzeros in the light-row
// kernel, 64 in the segment kernel) with the index loads two blocks and the
// factor gathers one block ahead of the MFMAs (GramStream).
// The 32x32 accumulator fragment (lane l, register r holds
// G[8(r/4) + 4(l/32) + r%4][l%32]) of two rows is exchanged between the wave
// halves so that lane i of half h owns row i of P of row h; the Cholesky with
// the forward solve fused in and the back solve then run per half with the
// current column broadcast through LDS, using reciprocal / reciprocal-sqrt
// multiplies.
// Adaptive noise: for the last mode the train SSE of the new samples is
// accumulated as u^T G u - 2 u^T c + sum r^2 per row, with
// u^T G u = (|L^{-1} b + z|^2 - u^T Lambda u) / alpha (since L^T u = L^{-1} b + z),
// so no separate pass over the nonzeros is needed.
#include <stdexcept>

#include "batched.h"

namespace mgpu {

constexpr int kPartial = 1024 + 64;  // floats per segment partial: fragment + RHS lanes

int batched_supported(int KP, int nother) {
  return (KP == 32 || KP == 64 || KP == 128) && nother >= 1 && nother <= kMaxBatchedOther;
}

bool batched_fits(const FactorSet& f, int nother, int KP) {
  for (int t = 0; t < nother; t++)
    if ((int64_t)f.n[t] * KP * (int64_t)sizeof(float) >= ((int64_t)1 << 31)) return false;
  return true;
}

// ---------------------------------------------------------------------------
// Gram stream.  The nonzeros of one or two rows are consumed in blocks of
// BLK = 32 or 64 nonzeros (a block never spans two rows).  Per block:
//   1. each of the first BLK lanes loads one nonzero's indices and residual
//      (coalesced), two blocks ahead of the block being multiplied;
//   2. the block is staged in this wave's LDS, and each lane reads, per
//      16-nonzero chunk, the 8 indices of its half with two 16-byte broadcast
//      reads, then gathers component l%32 of those 8 factor columns with
//      bounds-checked buffer loads (index -1 = padding reads as 0), one block
//      ahead of the block being multiplied;
//   3. the block's BLK / 16 chunks go through the split-BF16 MFMA (a chunk
//      that is all padding is skipped).
// So each wave keeps the index loads of block k+2 and the BLK / 2 gathers of
// block k+1 in flight while it multiplies block k.
// ---------------------------------------------------------------------------
// Block sizes: 32 nonzeros in the light-row kernel (its values in flight
// take 16 registers per block, which keeps it at 4 waves per SIMD), 64 in the
// segment kernel (few registers; more loads in flight per wave).
constexpr int kBlkLight = 32, kBlkSeg = 64;

template <int NO, int BLK>
struct BlkLds {
  int ix[2][NO][BLK];  // double buffered
  float r[2][BLK];
};


template <int NO, int BLK>
__device__ __forceinline__ void stage_blk(BlkLds<NO, BLK>& L, int buf, int lane, const BlkReg<NO>& b) {
  if (BLK == 64 || lane < BLK) {
#pragma unroll
    for (int t = 0; t < NO; t++) L.ix[buf][t][lane] = b.ix[t];
    L.r[buf][lane] = b.r;
  }
}


template <int NO, int BLK>
__device__ __forceinline__ void gather_blk(const Factors<NO>& F, const BlkLds<NO, BLK>& L, int buf, int half,
                                           int comp, float v[BLK / 2]) {
#pragma unroll
  for (int c = 0; c < BLK / 16; c++) {
    const int o = c * 16 + half * 8;
    int ix[NO][8];
#pragma unroll
    for (int t = 0; t < NO; t++)
#pragma unroll
      for (int i = 0; i < 8; i++) ix[t][i] = L.ix[buf][t][o + i];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      float x = ld_factor<32>(F.rs[0], ix[0][i], comp);
#pragma unroll
      for (int t = 1; t < NO; t++) x *= ld_factor<32>(F.rs[t], ix[t][i], comp);
      v[c * 8 + i] = x;
    }
  }
}

template <int NO, int BLK>
__device__ __forceinline__ void mfma_blk(const BlkLds<NO, BLK>& L, int buf, int half, const float v[BLK / 2],
                                         int nvalid, f32x16& acc, float& bp) {
#pragma unroll
  for (int c = 0; c < BLK / 16; c++) {
    if (c * 16 >= nvalid) break;
    const int o = c * 16 + half * 8;
    bf16x8 hi, lo;
    split8(v + c * 8, hi, lo);
#pragma unroll
    for (int i = 0; i < 8; i++) bp += v[c * 8 + i] * L.r[buf][o + i];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, lo, acc, 0, 0, 0);
  }
}

// Gram and RHS partials of nonzeros [begA, endA) into accA/bpA and, when
// PAIR, of [begB, endB) into accB/bpB, as one pipelined block stream.
template <int NO, bool PAIR, int BLK>
struct GramStream {
  const SampleArgs& a;
  const Factors<NO>& F;
  BlkLds<NO, BLK>& L;
  int begA, endA, begB, endB, nbA, nb, lane, half, comp;
  BlkReg<NO> nxt;

  __device__ __forceinline__ GramStream(const SampleArgs& a_, const Factors<NO>& F_, BlkLds<NO, BLK>& L_, int bA,
                                        int eA, int bB, int eB)
      : a(a_), F(F_), L(L_), begA(bA), endA(eA), begB(bB), endB(eB) {
    lane = threadIdx.x & 63;
    half = lane >> 5;
    comp = lane & 31;
    nbA = (endA - begA + BLK - 1) / BLK;
    nb = nbA + (PAIR ? (endB - begB + BLK - 1) / BLK : 0);
  }
  __device__ __forceinline__ int start(int k) const { return k < nbA ? begA + k * BLK : begB + (k - nbA) * BLK; }
  __device__ __forceinline__ int stop(int k) const { return k < nbA ? endA : endB; }

  // multiply block k (gathered in vc) and gather block k + 1 into vn
  __device__ __forceinline__ void step(int k, const float vc[BLK / 2], float vn[BLK / 2], f32x16& accA,
                                       float& bpA, f32x16& accB, float& bpB) {
    BlkReg<NO> after;
    if (k + 2 < nb) load_blk<NO, BLK>(a, start(k + 2), stop(k + 2), lane, after);
    if (k + 1 < nb) {
      stage_blk<NO>(L, (k + 1) & 1, lane, nxt);
      lds_wait();
      gather_blk<NO>(F, L, (k + 1) & 1, half, comp, vn);
    }
    const int nvalid = stop(k) - start(k);
    if (PAIR && k >= nbA) mfma_blk<NO>(L, k & 1, half, vc, nvalid, accB, bpB);
    else mfma_blk<NO>(L, k & 1, half, vc, nvalid, accA, bpA);
    nxt = after;
  }

  // index/residual loads of the stream's first two blocks for the ranges
  // [r.x, r.y) and [r.z, r.w) (what start(0..1) / stop(0..1) address); a
  // caller issues them ahead of time and hands them to run()
  static __device__ __forceinline__ void first_blocks(const SampleArgs& a, int4 r, int lane, BlkReg<NO>& b0,
                                                      BlkReg<NO>& b1) {
    const int na = (r.y - r.x + BLK - 1) / BLK;
    load_blk<NO, BLK>(a, na > 0 ? r.x : r.z, na > 0 ? r.y : r.w, lane, b0);
    load_blk<NO, BLK>(a, na > 1 ? r.x + BLK : r.z + (1 - na) * BLK, na > 1 ? r.y : r.w, lane, b1);
  }

  __device__ __forceinline__ void run(const BlkReg<NO>& b0, const BlkReg<NO>& b1, f32x16& accA, float& bpA,
                                      f32x16& accB, float& bpB) {
    if (nb == 0) return;
    float v0[BLK / 2], v1[BLK / 2];
    nxt = b1;
    stage_blk<NO>(L, 0, lane, b0);
    lds_wait();
    gather_blk<NO>(F, L, 0, half, comp, v0);
    for (int k = 0; k < nb; k += 2) {
      step(k, v0, v1, accA, bpA, accB, bpB);
      if (k + 1 < nb) step(k + 1, v1, v0, accA, bpA, accB, bpB);
    }
  }
};

// L of each half as its lower triangle, row i at i (i + 1) / 2; half 1
// starts kTri32 floats in, a multiple of 32 banks away from half 0, so the
// two halves' reads of one row in the back solve hit disjoint banks
constexpr int kTri32 = 544;

struct PairLds {
  alignas(16) float WT[2][4][32];   // current block columns of -P, per half (column-major)
  alignas(16) float Wb[2][4][32];   // RHS and the three probe vectors, per half
  alignas(16) float Lf[2 * kTri32]; // L per half (lower triangle)
  alignas(16) float us[2][32];      // the new samples (SSE)
};

// Finish two rows whose Gram fragments and RHS partials are in accA/accB,
// bpA/bpB (rowB < 0: no second row).  Returns this lane's SSE share.
//
// P = Lambda + alpha G stays in the MFMA accumulator layout (lane l, register
// r holds entry (8(r/4) + 4(l/32) + r%4, l%32)), negated, for the whole
// factorisation.  Right-looking Cholesky in 4-column blocks, k0 = 0, 4, .., 28:
//   1. the lanes holding columns k0..k0+3 publish them to LDS; lane i of half
//      h (row i of the pair's row h) reads its 4 entries, the 4 x 4 diagonal
//      block and its RHS, factors the block redundantly (L_D, y_D), forms its
//      panel row p = A[i][k0..k0+3] L_D^{-T} -- for a lane inside the block
//      this recurrence yields its row of L_D, zero above the diagonal -- and
//      eliminates its RHS;
//   2. the rank-4 update of both matrices is four v_mfma_f32_32x32x2_f32 on
//      the panel rows of the lanes below the block, whose A/B operands are
//      formed by v_permlane32_swap (rows of matrix A live in lanes 0-31, of
//      matrix B in lanes 32-63).
// L rows go to LDS as they are produced, for the back solve.  With
// refinement on, three probe vectors x (the row's normals z, all ones,
// alternating signs) are forward-solved beside the RHS, giving |L^{-1} x|^2 =
// x^T P^{-1} x for the condition estimate (probe_estimate).
__device__ float finish_pair(const SampleArgs& a, int rowA, int rowB, f32x16 accA, f32x16 accB, float bpA,
                             float bpB, int it, float alpha, PairLds& s, int nnzA, int nnzB) {
  int lane = threadIdx.x & 63;
  opaque(lane);
  const int half = lane >> 5;
  int i = lane & 31;
  const int row = half ? rowB : rowA;
  const bool active = row >= 0;
  // -P = -(Lambda + alpha G) in fragment layout (Lambda_f is symmetric and
  // 4 consecutive registers hold 4 consecutive rows of one column)
  {
    const float* lc = a.Lambda_f + i * 32 + 4 * half;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const float4 l4 = *reinterpret_cast<const float4*>(lc + 8 * m);
      accA[4 * m] = -fmaf(alpha, accA[4 * m], l4.x);
      accA[4 * m + 1] = -fmaf(alpha, accA[4 * m + 1], l4.y);
      accA[4 * m + 2] = -fmaf(alpha, accA[4 * m + 2], l4.z);
      accA[4 * m + 3] = -fmaf(alpha, accA[4 * m + 3], l4.w);
      accB[4 * m] = -fmaf(alpha, accB[4 * m], l4.x);
      accB[4 * m + 1] = -fmaf(alpha, accB[4 * m + 1], l4.y);
      accB[4 * m + 2] = -fmaf(alpha, accB[4 * m + 2], l4.z);
      accB[4 * m + 3] = -fmaf(alpha, accB[4 * m + 3], l4.w);
    }
  }
  const float recvb = __shfl_xor(half ? bpA : bpB, 32, 64);
  const float c = (half ? bpB : bpA) + recvb;
  float b = alpha * c;
  if (a.Lmu_rows) {
    if (active) b += a.Lmu_rows[(size_t)row * 32 + i];
  } else {
    b += a.Lmu_f[i];
  }
  bool bad = false;
  float inv_diag = 1.f;  // 1 / L[i][i]
  const float zi = i < a.K ? normal_f32(a.seed, (uint32_t)(active ? row : 0), (uint32_t)i, (uint32_t)it,
                                        make_tag(TAG_LATENT, a.mode))
                           : 0.f;
  // condition probes (refinement only): L^{-1} D^{1/2} x for x = z, ones,
  // alternating signs, D = diag(P)
  const bool probe = a.refine.count != nullptr && (nnzA >= a.refine.min_nnz || nnzB >= a.refine.min_nnz);
  float x1 = zi, x2 = i < a.K ? 1.f : 0.f, x3 = i < a.K ? ((i & 1) ? -1.f : 1.f) : 0.f;
  if (probe) {
    const float dA = frag_diag(accA, lane), dB = frag_diag(accB, lane);
    const float sd = sqrtf(-(half ? dB : dA));  // the fragments hold -P
    x1 *= sd;
    x2 *= sd;
    x3 *= sd;
  }
  const int tri_i = i * (i + 1) / 2;
  float* const Lh = s.Lf + half * kTri32;
#pragma unroll
  for (int k0 = 0; k0 < 32; k0 += 4) {
    opaque(i);
    // 1. publish columns k0..k0+3 of -P (lanes k0..k0+3 of both halves)
    const int m = k0 >> 3;
    if ((unsigned)(i - k0) < 4u) {
#pragma unroll
      for (int mm = 0; mm < 4; mm++) {
        if (mm < m) continue;  // rows above the block row are not read
        *reinterpret_cast<float4*>(&s.WT[0][i - k0][8 * mm + 4 * half]) =
            make_float4(accA[4 * mm], accA[4 * mm + 1], accA[4 * mm + 2], accA[4 * mm + 3]);
        *reinterpret_cast<float4*>(&s.WT[1][i - k0][8 * mm + 4 * half]) =
            make_float4(accB[4 * mm], accB[4 * mm + 1], accB[4 * mm + 2], accB[4 * mm + 3]);
      }
    }
    s.Wb[half][0][i] = b;
    if (probe) {
      s.Wb[half][1][i] = x1;
      s.Wb[half][2][i] = x2;
      s.Wb[half][3][i] = x3;
    }
    lds_wait();
    const float w0 = -s.WT[half][0][i], w1 = -s.WT[half][1][i], w2 = -s.WT[half][2][i], w3 = -s.WT[half][3][i];
    const float4 c0 = *reinterpret_cast<const float4*>(&s.WT[half][0][k0]);  // -D[.][0]
    const float4 c1 = *reinterpret_cast<const float4*>(&s.WT[half][1][k0]);
    const float4 c2 = *reinterpret_cast<const float4*>(&s.WT[half][2][k0]);
    const float4 c3 = *reinterpret_cast<const float4*>(&s.WT[half][3][k0]);
    const float4 bd = *reinterpret_cast<const float4*>(&s.Wb[half][0][k0]);
    // L_D = chol(D), D[t][u] = -c_u.(t), lower entries
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
    // panel row of this lane
    const float p0 = w0 * r0;
    const float p1 = (w1 - p0 * l10) * r1;
    const float p2 = (w2 - p0 * l20 - p1 * l21) * r2;
    const float p3 = (w3 - p0 * l30 - p1 * l31 - p2 * l32) * r3;
    if (k0 <= i) {
      float* lr = Lh + tri_i + k0;
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
      x1 = fwd(x1, *reinterpret_cast<const float4*>(&s.Wb[half][1][k0]));
      x2 = fwd(x2, *reinterpret_cast<const float4*>(&s.Wb[half][2][k0]));
      x3 = fwd(x3, *reinterpret_cast<const float4*>(&s.Wb[half][3][k0]));
    }
    // 2. -P += p p^T over the rows below the block, both matrices
    if (k0 + 4 < 32) {
      const float q0 = below ? p0 : 0.f, q1 = below ? p1 : 0.f, q2 = below ? p2 : 0.f, q3 = below ? p3 : 0.f;
      // swap(old, src): old's lanes 32-63 <-> src's lanes 0-31, so
      // .x = (q0 of A rows | q1 of A rows), .y = (q0 of B rows | q1 of B rows)
      const auto s01 = __builtin_amdgcn_permlane32_swap(__float_as_uint(q0), __float_as_uint(q1), false, false);
      const auto s23 = __builtin_amdgcn_permlane32_swap(__float_as_uint(q2), __float_as_uint(q3), false, false);
      const float a01 = __uint_as_float(s01[0]), b01 = __uint_as_float(s01[1]);
      const float a23 = __uint_as_float(s23[0]), b23 = __uint_as_float(s23[1]);
      accA = __builtin_amdgcn_mfma_f32_32x32x2f32(a01, a01, accA, 0, 0, 0);
      accA = __builtin_amdgcn_mfma_f32_32x32x2f32(a23, a23, accA, 0, 0, 0);
      accB = __builtin_amdgcn_mfma_f32_32x32x2f32(b01, b01, accB, 0, 0, 0);
      accB = __builtin_amdgcn_mfma_f32_32x32x2f32(b23, b23, accB, 0, 0, 0);
    }
  }
  if (bad && active && i == 0) atomicMax(a.fail, row + 1);
  b += zi;
  const float w = b;  // L^{-1} b + z
  // back solve L^T u = w.  Lane i needs column i of L (L[j][i], j > i),
  // i.e. entries of other lanes' rows: it reads them from the rows in LDS,
  // eight steps' worth at a time; the solved values are broadcast with
  // readlane.  (For lanes i > j the read lands in later rows and is unused.)
  lds_wait();
#pragma unroll
  for (int g = 31; g >= 0; g -= 8) {
    opaque(i);
    float Lc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) Lc[u] = Lh[(g - u) * (g - u + 1) / 2 + i];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int j = g - u;
      const float xl = b * inv_diag;  // final value on lane j
      const float xa = rdlane(xl, j), xb = rdlane(xl, 32 + j);
      const float xj = half ? xb : xa;
      b = (i < j) ? b - Lc[u] * xj : (i == j ? xj : b);
    }
  }
  if (active) a.U[(size_t)row * 32 + i] = b;
  if (probe) {
    const float est = probe_estimate<32>(a.K, x1, x2, x3, zi);
    refine_enqueue<32, 32>(a, active ? row : -1, half ? nnzB : nnzA, i, half * 32, est, Lh + tri_i);
  }
  if (!a.sse_partial) return 0.f;
  // train SSE share of this lane: (w^2 - u (Lambda u)) / alpha - 2 u c (+ sum r^2)
  s.us[half][i] = b;
  lds_wait();
  const float* Lrow = a.Lambda_f + i * 32;
  float lu = 0.f;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const float4 u4 = *reinterpret_cast<const float4*>(&s.us[half][4 * q]);
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


// Each wave takes pairs q, q + G, ... (G = grid).  The pair's row ids and nnz
// ranges come from the plan in one load, and the next pair's ranges and its
// first two index blocks are loaded while the current pair is multiplied and
// finished, so a pair's stream starts with its indices already in registers.
template <int NO>
__global__ void __launch_bounds__(64, 4) k_light32(SampleArgs a, BatchPlan p) {
  __shared__ PairLds s;
  __shared__ BlkLds<NO, kBlkLight> blk;
  const Factors<NO> F = make_factors<NO, 32>(a.others);
  const int it = *a.iter;
  const float alpha = (float)(*a.alpha);
  const int lane = threadIdx.x;
  double sse = 0.0;
  int q = blockIdx.x;
  int4 rng = make_int4(0, 0, 0, 0);
  int2 rows = make_int2(-1, -1);
  BlkReg<NO> b0, b1;
  if (q < p.n_pair) {
    rng = p.pair_rng[q];
    rows = p.pair_row[q];
    GramStream<NO, true, kBlkLight>::first_blocks(a, rng, lane, b0, b1);
  }
  for (; q < p.n_pair; q += gridDim.x) {
    const int qn = q + gridDim.x;
    int4 rng_n = make_int4(0, 0, 0, 0);
    int2 rows_n = make_int2(-1, -1);
    if (qn < p.n_pair) {
      rng_n = p.pair_rng[qn];
      rows_n = p.pair_row[qn];
    }
    f32x16 accA, accB;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      accA[r] = 0.f;
      accB[r] = 0.f;
    }
    float bpA = 0.f, bpB = 0.f;
    GramStream<NO, true, kBlkLight>(a, F, blk, rng.x, rng.y, rng.z, rng.w).run(b0, b1, accA, bpA, accB, bpB);
    if (qn < p.n_pair) GramStream<NO, true, kBlkLight>::first_blocks(a, rng_n, lane, b0, b1);
    sse += (double)finish_pair(a, rows.x, rows.y, accA, accB, bpA, bpB, it, alpha, s, rng.y - rng.x,
                               rng.w - rng.z);
    rng = rng_n;
    rows = rows_n;
  }
  write_sse(a, sse, blockIdx.x);
}

template <int NO>
__global__ void __launch_bounds__(64) k_seg32(SampleArgs a, BatchPlan p) {
  __shared__ BlkLds<NO, kBlkSeg> blk;
  const Factors<NO> F = make_factors<NO, 32>(a.others);
  const int lane = threadIdx.x;
  for (int sg = blockIdx.x; sg < p.n_seg; sg += gridDim.x) {
    f32x16 acc, unused;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    float bp = 0.f, unused_bp = 0.f;
    const int4 r = make_int4(p.seg_beg[sg], p.seg_end[sg], 0, 0);
    BlkReg<NO> b0, b1;
    GramStream<NO, false, kBlkSeg>::first_blocks(a, r, lane, b0, b1);
    GramStream<NO, false, kBlkSeg>(a, F, blk, r.x, r.y, 0, 0).run(b0, b1, acc, bp, unused, unused_bp);
    float* out = p.partial + (size_t)sg * kPartial;
#pragma unroll
    for (int r = 0; r < 16; r++) out[r * 64 + lane] = acc[r];
    out[1024 + lane] = bp;
  }
}

// the summed partial of one heavy slot (k_seg_sum leaves it in the slot's
// first segment)
__device__ __forceinline__ void slot_partial(const BatchPlan& p, int slot, int lane, f32x16& acc, float& bp) {
  const float* in = p.partial + (size_t)p.heavy_seg0[slot] * kPartial;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = in[r * 64 + lane];
  bp = in[1024 + lane];
}

__global__ void __launch_bounds__(64, 4) k_heavy32(SampleArgs a, BatchPlan p, int sse_slot0) {
  __shared__ PairLds s;
  const int lane = threadIdx.x;
  const int it = *a.iter;
  const float alpha = (float)(*a.alpha);
  double sse = 0.0;
  const int npairs = (p.n_heavy + 1) / 2;
  for (int q = blockIdx.x; q < npairs; q += gridDim.x) {
    f32x16 accA, accB;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      accA[r] = 0.f;
      accB[r] = 0.f;
    }
    float bpA = 0.f, bpB = 0.f;
    const int rowA = p.heavy[2 * q];
    const int rowB = (2 * q + 1 < p.n_heavy) ? p.heavy[2 * q + 1] : -1;
    slot_partial(p, 2 * q, lane, accA, bpA);
    if (rowB >= 0) slot_partial(p, 2 * q + 1, lane, accB, bpB);
    sse += (double)finish_pair(a, rowA, rowB, accA, accB, bpA, bpB, it, alpha, s, p.seg_len, p.seg_len);
  }
  write_sse(a, sse, sse_slot0 + blockIdx.x);
}

int batched_light_grid(const BatchPlan& p) {
  if (p.KP >= 64) return batched64_light_grid(p);  // one row per work-group
  const int cap = p.light_cap > 0 ? p.light_cap : kDefaultLightCap;
  return p.n_pair < cap ? p.n_pair : cap;
}
int batched_heavy_grid(const BatchPlan& p) {
  if (p.KP >= 64) return batched64_heavy_grid(p);
  const int npairs = (p.n_heavy + 1) / 2;
  const int cap = p.heavy_cap > 0 ? p.heavy_cap : kDefaultHeavyCap;
  return npairs < cap ? npairs : cap;
}
int batched_seg_grid(const BatchPlan& p) {
  const int cap = p.seg_cap > 0 ? p.seg_cap : kDefaultSegCap;
  return p.n_seg < cap ? p.n_seg : cap;
}

template <int NO>
static void launch_batched_t(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg) {
  const int gl = batched_light_grid(p), gh = batched_heavy_grid(p);
  const bool fork = p.n_seg > 0 && seg.stream;
  if (p.n_seg > 0) {
    const hipStream_t ss = fork ? seg.stream : s;
    if (fork) seg_fork(s, seg);
    const int gs = batched_seg_grid(p);
    hipLaunchKernelGGL((k_seg32<NO>), dim3(gs), dim3(64), 0, ss, a, p);
    launch_seg_sum<kPartial>(p, ss);
    // the heavy rows depend only on their segments: finished on the segment
    // stream too, beside the light rows
    if (gh > 0) hipLaunchKernelGGL(k_heavy32, dim3(gh), dim3(64), 0, ss, a, p, gl);
    if (fork) seg_record_join(seg);
  }
  if (gl > 0) hipLaunchKernelGGL((k_light32<NO>), dim3(gl), dim3(64), 0, s, a, p);
  if (fork) seg_wait_join(s, seg);
  launch_refine(a, s);
}

void launch_sample_batched(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg) {
  if (!batched_fits(a.others, a.rows.nother, a.KP))
    throw std::runtime_error("batched sampler: a factor matrix exceeds 2^31 bytes");
  if (a.KP != p.KP || !batched_supported(a.KP, a.rows.nother))
    throw std::runtime_error("batched sampler: plan and latent size disagree or are unsupported");
  if (a.KP == 64) return launch_sample_batched64(a, p, s, seg);
  if (a.KP == 128) return launch_sample_batched128(a, p, s, seg);
  switch (a.rows.nother) {
    case 1: return launch_batched_t<1>(a, p, s, seg);
    case 2: return launch_batched_t<2>(a, p, s, seg);
    case 3: return launch_batched_t<3>(a, p, s, seg);
    case 4: return launch_batched_t<4>(a, p, s, seg);
    default: break;
  }
}

}  // namespace mgpu
