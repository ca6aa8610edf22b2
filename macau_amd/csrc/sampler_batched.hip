// sampler_batched.hip -- nnz-bucketed per-row sampler with a split-BF16
// MFMA Gram (K in (16, 32]).
//
// Same mathematics as sample_latent_blas (latentprior.cpp:361-385) and
// sample_latent_tensor (latentprior.cpp:118-161): per row
//   P = Lambda + alpha G,  G = sum_e c_e c_e^T,  b = alpha sum_e c_e r_e + Lambda mu_n,
//   u = L^{-T}(L^{-1} b + z),  P = L L^T.
//
// Work decomposition of one half-sweep (BatchPlan, built on the host):
//   * rows with more than 1024 nonzeros are split into 1024-nnz segments;
//     k_seg32 computes each segment's partial Gram and RHS (one wave per
//     segment), k_heavy32 sums the partials of two rows in segment order and
//     finishes them;
//   * the other rows, sorted by nnz (descending), are taken two per wave by
//     k_light32 (similar nnz in a pair keeps both halves busy).
// Gram: each MFMA step takes 16 nonzeros (lanes 0-31 the first 8, lanes
// 32-63 the next 8; lane l holds component l%32), splits every value into
// hi = bf16(x) and lo = bf16(x - hi), and accumulates G += Hi Hi^T + Hi Lo^T
// + Lo Hi^T with three v_mfma_f32_32x32x16_bf16 into FP32 (the dropped Lo Lo^T
// term is ~2^-18 relative).  The RHS accumulates in FP32 on the vector ALU.
// The nonzeros stream through in 64-nonzero blocks with the index loads two
// blocks and the factor gathers one block ahead of the MFMAs (GramStream).
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

#include "common.h"
#include "kernels.h"

namespace mgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kPartial = 1024 + 64;  // floats per segment partial: fragment + RHS lanes

int batched_supported(int KP, int nother) { return KP == 32 && nother >= 1 && nother <= kMaxOther; }

bool batched_fits(const FactorSet& f, int nother, int KP) {
  for (int t = 0; t < nother; t++)
    if ((int64_t)f.n[t] * KP * (int64_t)sizeof(float) >= ((int64_t)1 << 31)) return false;
  return true;
}

// ---------------------------------------------------------------------------
// Gram stream.  The nonzeros of one or two rows are consumed in 64-nonzero
// blocks (a block never spans two rows).  Per block:
//   1. every lane loads one nonzero's indices and residual (coalesced),
//      two blocks ahead of the block being multiplied;
//   2. the block is staged in this wave's LDS, and each lane reads, per
//      16-nonzero chunk, the 8 indices of its half with two 16-byte broadcast
//      reads, then gathers component l%32 of those 8 factor columns with
//      bounds-checked buffer loads (index -1 = padding reads as 0), one block
//      ahead of the block being multiplied;
//   3. the block's four chunks go through the split-BF16 MFMA (a chunk that
//      is all padding is skipped).
// So each wave keeps the index loads of block k+2 and the 32 gathers of block
// k+1 in flight while it multiplies block k.
// ---------------------------------------------------------------------------
constexpr int kBlk = 64;

template <int NO>
struct BlkLds {
  int ix[2][NO][kBlk];  // double buffered
  float r[2][kBlk];
};

template <int NO>
struct BlkReg {
  int ix[NO];
  float r;
};

template <int NO>
struct Factors {
  __amdgpu_buffer_rsrc_t rs[NO];
};

// buffer descriptors of the other modes' factors (KP = 32 floats per column);
// word 3 = 0x00020000 selects 32-bit data, raw (unswizzled) addressing
template <int NO>
__device__ __forceinline__ Factors<NO> make_factors(const FactorSet& f) {
  Factors<NO> r;
#pragma unroll
  for (int t = 0; t < NO; t++)
    r.rs[t] = __builtin_amdgcn_make_buffer_rsrc((void*)f.f[t], (short)0, f.n[t] * 32 * (int)sizeof(float),
                                                0x00020000);
  return r;
}

template <int NO>
__device__ __forceinline__ void load_blk(const SampleArgs& a, int e0, int end, int lane, BlkReg<NO>& b) {
  const int e = e0 + lane;
  if (e < end) {
#pragma unroll
    for (int t = 0; t < NO; t++) b.ix[t] = a.rows.idx[(size_t)e * NO + t];
    b.r = a.rows.r[e];
  } else {
#pragma unroll
    for (int t = 0; t < NO; t++) b.ix[t] = -1;
    b.r = 0.f;
  }
}

// LDS writes of this wave must land before its lanes read each other's
// entries: wait for the LDS queue (lgkmcnt(0), vmcnt/expcnt untouched).
__device__ __forceinline__ void lds_wait() {
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
}

template <int NO>
__device__ __forceinline__ void stage_blk(BlkLds<NO>& L, int buf, int lane, const BlkReg<NO>& b) {
#pragma unroll
  for (int t = 0; t < NO; t++) L.ix[buf][t][lane] = b.ix[t];
  L.r[buf][lane] = b.r;
}

__device__ __forceinline__ float ld_factor(__amdgpu_buffer_rsrc_t rs, int ix, int comp) {
  // ix = -1 gives an offset past num_records: the buffer load returns 0
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, ix * 128 + comp * 4, 0, 0));
}

template <int NO>
__device__ __forceinline__ void gather_blk(const Factors<NO>& F, const BlkLds<NO>& L, int buf, int half, int comp,
                                           float v[32]) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int o = c * 16 + half * 8;
    int ix[NO][8];
#pragma unroll
    for (int t = 0; t < NO; t++)
#pragma unroll
      for (int i = 0; i < 8; i++) ix[t][i] = L.ix[buf][t][o + i];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      float x = ld_factor(F.rs[0], ix[0][i], comp);
#pragma unroll
      for (int t = 1; t < NO; t++) x *= ld_factor(F.rs[t], ix[t][i], comp);
      v[c * 8 + i] = x;
    }
  }
}

template <int NO>
__device__ __forceinline__ void mfma_blk(const BlkLds<NO>& L, int buf, int half, const float v[32], int nvalid,
                                         f32x16& acc, float& bp) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (c * 16 >= nvalid) break;
    const int o = c * 16 + half * 8;
    bf16x8 hi, lo;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float x = v[c * 8 + i];
      const __bf16 h = (__bf16)x;
      hi[i] = h;
      lo[i] = (__bf16)(x - (float)h);
      bp += x * L.r[buf][o + i];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, hi, acc, 0, 0, 0);
  }
}

// Gram and RHS partials of nonzeros [begA, endA) into accA/bpA and, when
// PAIR, of [begB, endB) into accB/bpB, as one pipelined block stream.
template <int NO, bool PAIR>
struct GramStream {
  const SampleArgs& a;
  const Factors<NO>& F;
  BlkLds<NO>& L;
  int begA, endA, begB, endB, nbA, nb, lane, half, comp;
  BlkReg<NO> nxt;

  __device__ __forceinline__ GramStream(const SampleArgs& a_, const Factors<NO>& F_, BlkLds<NO>& L_, int bA,
                                        int eA, int bB, int eB)
      : a(a_), F(F_), L(L_), begA(bA), endA(eA), begB(bB), endB(eB) {
    lane = threadIdx.x & 63;
    half = lane >> 5;
    comp = lane & 31;
    nbA = (endA - begA + kBlk - 1) / kBlk;
    nb = nbA + (PAIR ? (endB - begB + kBlk - 1) / kBlk : 0);
  }
  __device__ __forceinline__ int start(int k) const { return k < nbA ? begA + k * kBlk : begB + (k - nbA) * kBlk; }
  __device__ __forceinline__ int stop(int k) const { return k < nbA ? endA : endB; }

  // multiply block k (gathered in vc) and gather block k + 1 into vn
  __device__ __forceinline__ void step(int k, const float vc[32], float vn[32], f32x16& accA, float& bpA,
                                       f32x16& accB, float& bpB) {
    BlkReg<NO> after;
    if (k + 2 < nb) load_blk<NO>(a, start(k + 2), stop(k + 2), lane, after);
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

  __device__ __forceinline__ void run(f32x16& accA, float& bpA, f32x16& accB, float& bpB) {
    if (nb == 0) return;
    float v0[32], v1[32];
    BlkReg<NO> b0;
    load_blk<NO>(a, start(0), stop(0), lane, b0);
    if (nb > 1) load_blk<NO>(a, start(1), stop(1), lane, nxt);
    stage_blk<NO>(L, 0, lane, b0);
    lds_wait();
    gather_blk<NO>(F, L, 0, half, comp, v0);
    for (int k = 0; k < nb; k += 2) {
      step(k, v0, v1, accA, bpA, accB, bpB);
      if (k + 1 < nb) step(k + 1, v1, v0, accA, bpA, accB, bpB);
    }
  }
};

struct PairLds {
  alignas(16) float colb[2][2][32];  // column broadcast, double buffered, per half
  alignas(16) float bcol[2][2][32];  // RHS broadcast, double buffered, per half
  alignas(16) float Lf[2][32][36];   // L per half, row i at Lf[half][i]
  alignas(16) float us[2][32];       // the new samples (SSE)
};

__device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// An empty asm that "changes" v: predicates built from v after it are formed
// at that point instead of being hoisted and held in scalar registers.
__device__ __forceinline__ void opaque(int& v) { asm volatile("" : "+v"(v)); }

// Finish two rows whose Gram fragments and RHS partials are in accA/accB,
// bpA/bpB (rowB < 0: no second row).  Lane i of half h owns row i of P for
// row h.  Returns this lane's SSE share.
__device__ float finish_pair(const SampleArgs& a, int rowA, int rowB, const f32x16& accA, const f32x16& accB,
                             float bpA, float bpB, int it, PairLds& s) {
  int lane = threadIdx.x & 63;
  opaque(lane);
  const int half = lane >> 5;
  int i = lane & 31;
  const int row = half ? rowB : rowA;
  const bool active = row >= 0;
  const float alpha = (float)(*a.alpha);
  // fragments -> row i of this half's P (G is symmetric: column i = row i)
  float P[32];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const float send = half ? accA[r] : accB[r];
    const float recv = __shfl_xor(send, 32, 64);
    const float own = half ? accB[r] : accA[r];
    const int j0 = 8 * (r >> 2) + (r & 3);
    P[j0] = half ? recv : own;
    P[j0 + 4] = half ? own : recv;
  }
  const float recvb = __shfl_xor(half ? bpA : bpB, 32, 64);
  const float c = (half ? bpB : bpA) + recvb;
  const float* Lrow = a.Lambda_f + i * 32;
#pragma unroll
  for (int j = 0; j < 32; j++) P[j] = Lrow[j] + alpha * P[j];
  float b = alpha * c;
  if (a.Lmu_rows) {
    if (active) b += a.Lmu_rows[(size_t)row * 32 + i];
  } else {
    b += a.Lmu_f[i];
  }
  // Right-looking Cholesky with the forward solve fused in, branch-free.
  // Step k reads column k (= row k, P is symmetric) and the RHS of lane k
  // from LDS; lanes i > k eliminate (lanes i < k hold only upper-triangle
  // values that are never read again).  Column k + 1 is updated first and
  // published for the next step before the rest of the row, so its LDS
  // round trip overlaps this step's remaining FMAs.
  bool bad = false;
  float inv_diag = 1.f;  // 1 / L[i][i]
  s.colb[0][half][i] = P[0];
  s.bcol[0][half][i] = b;
#pragma unroll
  for (int k = 0; k < 32; k++) {
    opaque(i);
    lds_wait();
    const float* cb = s.colb[k & 1][half];
    float col[32];
#pragma unroll
    for (int q = k / 4; q < 8; q++) {
      const float4 t4 = *reinterpret_cast<const float4*>(cb + 4 * q);
      col[4 * q] = t4.x;
      col[4 * q + 1] = t4.y;
      col[4 * q + 2] = t4.z;
      col[4 * q + 3] = t4.w;
    }
    const float bk = s.bcol[k & 1][half][k];
    const bool badp = !(col[k] > 0.f);
    bad = bad || badp;
    const float rs = __builtin_amdgcn_rsqf(badp ? 1.f : col[k]);  // 1 / sqrt(d)
    const bool gt = i > k, eq = i == k;
    const float sc = gt ? P[k] * (rs * rs) : 0.f;
    const float bn = eq ? bk * rs : b - sc * bk;
    if (k + 1 < 32) {
      P[k + 1] -= sc * col[k + 1];
      s.colb[(k + 1) & 1][half][i] = P[k + 1];
      s.bcol[(k + 1) & 1][half][i] = bn;
    }
#pragma unroll
    for (int j = k + 2; j < 32; j++) P[j] -= sc * col[j];
    P[k] *= rs;  // L[i][k] for i >= k
    inv_diag = eq ? rs : inv_diag;
    b = bn;
  }
  if (bad && active && i == 0) atomicMax(a.fail, row + 1);
  if (i < a.K) b += normal_f32(a.seed, (uint32_t)(active ? row : 0), (uint32_t)i, (uint32_t)it,
                               make_tag(TAG_LATENT, a.mode));
  const float w = b;  // L^{-1} b + z
  // back solve L^T u = w.  Lane i needs column i of L (L[j][i], j > i),
  // which is row data of other lanes: rows go to LDS and each lane reads
  // its column once; the solved values are broadcast with readlane.
  float* Lr = s.Lf[half][i];
#pragma unroll
  for (int q = 0; q < 8; q++)
    *reinterpret_cast<float4*>(Lr + 4 * q) = make_float4(P[4 * q], P[4 * q + 1], P[4 * q + 2], P[4 * q + 3]);
  lds_wait();
  float Lc[32];
#pragma unroll
  for (int j = 0; j < 32; j++) Lc[j] = s.Lf[half][j][i];
#pragma unroll
  for (int j = 31; j >= 0; j--) {
    opaque(i);
    const float xl = b * inv_diag;  // final value on lane j
    const float xj = half ? rdlane(xl, 32 + j) : rdlane(xl, j);
    b = (i < j) ? b - Lc[j] * xj : (i == j ? xj : b);
  }
  if (active) a.U[(size_t)row * 32 + i] = b;
  if (!a.sse_partial) return 0.f;
  // train SSE share of this lane: (w^2 - u (Lambda u)) / alpha - 2 u c (+ sum r^2)
  s.us[half][i] = b;
  lds_wait();
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

__device__ inline void write_sse(const SampleArgs& a, double acc, int slot) {
  if (!a.sse_partial) return;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x == 0) a.sse_partial[slot] = acc;
}

template <int NO>
__global__ void __launch_bounds__(64, 3) k_light32(SampleArgs a, BatchPlan p) {
  __shared__ PairLds s;
  __shared__ BlkLds<NO> blk;
  const Factors<NO> F = make_factors<NO>(a.others);
  const int it = *a.iter;
  double sse = 0.0;
  const int npairs = (p.n_light + 1) / 2;
  for (int q = blockIdx.x; q < npairs; q += gridDim.x) {
    const int rowA = p.light[2 * q];
    const int rowB = (2 * q + 1 < p.n_light) ? p.light[2 * q + 1] : -1;
    const int begA = a.rows.row_ptr[rowA], endA = a.rows.row_ptr[rowA + 1];
    int begB = 0, endB = 0;
    if (rowB >= 0) {
      begB = a.rows.row_ptr[rowB];
      endB = a.rows.row_ptr[rowB + 1];
    }
    f32x16 accA, accB;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      accA[r] = 0.f;
      accB[r] = 0.f;
    }
    float bpA = 0.f, bpB = 0.f;
    GramStream<NO, true>(a, F, blk, begA, endA, begB, endB).run(accA, bpA, accB, bpB);
    sse += (double)finish_pair(a, rowA, rowB, accA, accB, bpA, bpB, it, s);
  }
  write_sse(a, sse, blockIdx.x);
}

template <int NO>
__global__ void __launch_bounds__(64) k_seg32(SampleArgs a, BatchPlan p) {
  __shared__ BlkLds<NO> blk;
  const Factors<NO> F = make_factors<NO>(a.others);
  const int lane = threadIdx.x;
  for (int sg = blockIdx.x; sg < p.n_seg; sg += gridDim.x) {
    f32x16 acc, unused;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    float bp = 0.f, unused_bp = 0.f;
    GramStream<NO, false>(a, F, blk, p.seg_beg[sg], p.seg_end[sg], 0, 0).run(acc, bp, unused, unused_bp);
    float* out = p.partial + (size_t)sg * kPartial;
#pragma unroll
    for (int r = 0; r < 16; r++) out[r * 64 + lane] = acc[r];
    out[1024 + lane] = bp;
  }
}

// sum of the segment partials of one heavy slot (fixed segment order)
__device__ __forceinline__ void sum_segments(const BatchPlan& p, int slot, int lane, f32x16& acc, float& bp) {
  for (int sg = p.heavy_seg0[slot]; sg < p.heavy_seg0[slot + 1]; sg++) {
    const float* in = p.partial + (size_t)sg * kPartial;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] += in[r * 64 + lane];
    bp += in[1024 + lane];
  }
}

__global__ void __launch_bounds__(64, 4) k_heavy32(SampleArgs a, BatchPlan p, int sse_slot0) {
  __shared__ PairLds s;
  const int lane = threadIdx.x;
  const int it = *a.iter;
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
    sum_segments(p, 2 * q, lane, accA, bpA);
    if (rowB >= 0) sum_segments(p, 2 * q + 1, lane, accB, bpB);
    sse += (double)finish_pair(a, rowA, rowB, accA, accB, bpA, bpB, it, s);
  }
  write_sse(a, sse, sse_slot0 + blockIdx.x);
}

int batched_light_grid(const BatchPlan& p) {
  const int npairs = (p.n_light + 1) / 2;
  return npairs < 8192 ? npairs : 8192;
}
int batched_heavy_grid(const BatchPlan& p) {
  const int npairs = (p.n_heavy + 1) / 2;
  return npairs < 2048 ? npairs : 2048;
}

template <int NO>
static void launch_batched_t(const SampleArgs& a, const BatchPlan& p, hipStream_t s) {
  const int gl = batched_light_grid(p), gh = batched_heavy_grid(p);
  if (p.n_seg > 0) {
    const int gs = p.n_seg < 16384 ? p.n_seg : 16384;
    hipLaunchKernelGGL((k_seg32<NO>), dim3(gs), dim3(64), 0, s, a, p);
  }
  if (gl > 0) hipLaunchKernelGGL((k_light32<NO>), dim3(gl), dim3(64), 0, s, a, p);
  if (gh > 0) hipLaunchKernelGGL(k_heavy32, dim3(gh), dim3(64), 0, s, a, p, gl);
}

void launch_sample_batched(const SampleArgs& a, const BatchPlan& p, hipStream_t s) {
  if (!batched_fits(a.others, a.rows.nother, a.KP))
    throw std::runtime_error("batched sampler: a factor matrix exceeds 2^31 bytes");
  switch (a.rows.nother) {
    case 1: return launch_batched_t<1>(a, p, s);
    case 2: return launch_batched_t<2>(a, p, s);
    case 3: return launch_batched_t<3>(a, p, s);
    case 4: return launch_batched_t<4>(a, p, s);
    default: break;
  }
}

}  // namespace mgpu
