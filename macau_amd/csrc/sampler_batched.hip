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
// Gram: each wave step gathers 16 nonzeros (lanes 0-31 the first 8, lanes
// 32-63 the next 8; lane l holds component l%32), splits every value into
// hi = bf16(x) and lo = bf16(x - hi), and accumulates G += Hi Hi^T + Hi Lo^T
// + Lo Hi^T with three v_mfma_f32_32x32x16_bf16 into FP32 (the dropped Lo Lo^T
// term is ~2^-18 relative).  The RHS accumulates in FP32 on the vector ALU.
// The gathers are software-pipelined: the indices and values of the next
// 16 nonzeros are in flight while the current factor columns load.
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
#include "common.h"
#include "kernels.h"

namespace mgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kPartial = 1024 + 64;  // floats per segment partial: fragment + RHS lanes

int batched_supported(int KP, int nother) { return KP == 32 && nother >= 1 && nother <= kMaxOther; }

// indices and residuals of the 8 nonzeros a lane gathers in one chunk
template <int NO>
struct Chunk {
  int ix[8][NO];
  float r[8];
};

template <int NO>
__device__ __forceinline__ void load_chunk(const SampleArgs& a, int e0, int end, int half, Chunk<NO>& c) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int e = e0 + half * 8 + i;
    if (e < end) {
#pragma unroll
      for (int t = 0; t < NO; t++) c.ix[i][t] = a.rows.idx[(size_t)e * NO + t];
      c.r[i] = a.rows.r[e];
    } else {
#pragma unroll
      for (int t = 0; t < NO; t++) c.ix[i][t] = -1;
      c.r[i] = 0.f;
    }
  }
}

template <int NO>
__device__ __forceinline__ void gather_chunk(const SampleArgs& a, const Chunk<NO>& c, int comp, float v[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    float x = 0.f;
    if (c.ix[i][0] >= 0) {
      x = a.others.f[0][(size_t)c.ix[i][0] * 32 + comp];
#pragma unroll
      for (int t = 1; t < NO; t++) x *= a.others.f[t][(size_t)c.ix[i][t] * 32 + comp];
    }
    v[i] = x;
  }
}

__device__ __forceinline__ void mfma_chunk(const float v[8], const float r[8], f32x16& acc, float& bp) {
  bf16x8 hi, lo;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const __bf16 h = (__bf16)v[i];
    hi[i] = h;
    lo[i] = (__bf16)(v[i] - (float)h);
    bp += v[i] * r[i];
  }
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo, hi, acc, 0, 0, 0);
}

// Gram and RHS partial of nonzeros [beg, end), pipelined one chunk ahead
template <int NO>
__device__ __forceinline__ void gram32(const SampleArgs& a, int beg, int end, int half, int comp, f32x16& acc,
                                       float& bp) {
  if (beg >= end) return;
  Chunk<NO> cur;
  load_chunk<NO>(a, beg, end, half, cur);
  for (int e0 = beg; e0 < end; e0 += 16) {
    float v[8];
    gather_chunk<NO>(a, cur, comp, v);
    Chunk<NO> nxt;
    if (e0 + 16 < end) load_chunk<NO>(a, e0 + 16, end, half, nxt);
    mfma_chunk(v, cur.r, acc, bp);
    cur = nxt;
  }
}

// Gram of two rows with their chunks interleaved (two independent chains)
template <int NO>
__device__ __forceinline__ void gram32_pair(const SampleArgs& a, int begA, int endA, int begB, int endB, int half,
                                            int comp, f32x16& accA, float& bpA, f32x16& accB, float& bpB) {
  const int nA = endA - begA, nB = endB - begB;
  const int nmin = nA < nB ? nA : nB;
  if (nmin > 0) {
    Chunk<NO> ca, cb;
    load_chunk<NO>(a, begA, endA, half, ca);
    load_chunk<NO>(a, begB, endB, half, cb);
    for (int o = 0; o < nmin; o += 16) {
      float va[8], vb[8];
      gather_chunk<NO>(a, ca, comp, va);
      gather_chunk<NO>(a, cb, comp, vb);
      Chunk<NO> na, nb;
      if (o + 16 < nA) load_chunk<NO>(a, begA + o + 16, endA, half, na);
      if (o + 16 < nB) load_chunk<NO>(a, begB + o + 16, endB, half, nb);
      mfma_chunk(va, ca.r, accA, bpA);
      mfma_chunk(vb, cb.r, accB, bpB);
      ca = na;
      cb = nb;
    }
  }
  const int done = (nmin + 15) & ~15;
  if (done < nA) gram32<NO>(a, begA + done, endA, half, comp, accA, bpA);
  if (done < nB) gram32<NO>(a, begB + done, endB, half, comp, accB, bpB);
}

struct PairLds {
  float colb[2][2][36];  // column broadcast, double buffered, per half (16-byte aligned rows)
  float Lp[2][528];      // packed lower L per half, row i at i(i+1)/2
  float xs[2][2];        // back-solve broadcast, double buffered, per half
  float us[2][32];       // the new samples (SSE)
};

// Finish two rows whose Gram fragments and RHS partials are in accA/accB,
// bpA/bpB (rowB < 0: no second row).  Returns this lane's SSE share.
__device__ float finish_pair(const SampleArgs& a, int rowA, int rowB, const f32x16& accA, const f32x16& accB,
                             float bpA, float bpB, int it, PairLds& s) {
  const int lane = threadIdx.x, half = lane >> 5, i = lane & 31;
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
  // Cholesky with fused forward solve
  bool bad = false;
  float inv_diag = 1.f;  // 1 / L[i][i]
#pragma unroll
  for (int k = 0; k < 32; k++) {
    float* cb = s.colb[k & 1][half];
    if (i >= k) cb[i] = P[k];
    if (i == k) cb[32] = b;
    __syncthreads();
    float dkk = cb[k];
    if (!(dkk > 0.f)) {
      bad = bad || (i == k);
      dkk = 1.f;
    }
    const float rs = __builtin_amdgcn_rsqf(dkk);  // 1 / sqrt(d)
    const float bk = cb[32];
    if (i > k) {
      const float sc = P[k] * (rs * rs);
#pragma unroll
      for (int j = k + 1; j < 32; j++) P[j] -= sc * cb[j];
      b -= sc * bk;
      P[k] = P[k] * rs;
    } else if (i == k) {
      P[k] = dkk * rs;
      inv_diag = rs;
      b = bk * rs;
    }
  }
  if (bad && active) atomicMax(a.fail, row + 1);
  if (i < a.K) b += normal_f32(a.seed, (uint32_t)(active ? row : 0), (uint32_t)i, (uint32_t)it,
                               make_tag(TAG_LATENT, a.mode));
  const float w = b;  // L^{-1} b + z
  // back solve L^T x = w with L packed in LDS
  float* Lp = s.Lp[half];
  const int ri = i * (i + 1) / 2;
#pragma unroll
  for (int k = 0; k < 32; k++)
    if (k <= i) Lp[ri + k] = P[k];
  __syncthreads();
#pragma unroll
  for (int j = 31; j >= 0; j--) {
    const float lji = (i <= j) ? Lp[j * (j + 1) / 2 + i] : 0.f;  // L[j][i]
    if (i == j) s.xs[j & 1][half] = b * inv_diag;
    __syncthreads();
    const float xj = s.xs[j & 1][half];
    if (i < j) b -= lji * xj;
    if (i == j) b = xj;
  }
  if (active) a.U[(size_t)row * 32 + i] = b;
  if (!a.sse_partial) return 0.f;
  // train SSE share of this lane: (w^2 - u (Lambda u)) / alpha - 2 u c (+ sum r^2)
  s.us[half][i] = b;
  __syncthreads();
  float lu = 0.f;
#pragma unroll
  for (int j = 0; j < 32; j++) lu += Lrow[j] * s.us[half][j];
  float t = 0.f;
  if (active) {
    t = (w * w - b * lu) / alpha - 2.f * b * c;
    if (i == 0) t += a.r2[row];
  }
  __syncthreads();
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
  const int lane = threadIdx.x, half = lane >> 5, comp = lane & 31;
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
    gram32_pair<NO>(a, begA, endA, begB, endB, half, comp, accA, bpA, accB, bpB);
    sse += (double)finish_pair(a, rowA, rowB, accA, accB, bpA, bpB, it, s);
  }
  write_sse(a, sse, blockIdx.x);
}

template <int NO>
__global__ void __launch_bounds__(64) k_seg32(SampleArgs a, BatchPlan p) {
  const int lane = threadIdx.x, half = lane >> 5, comp = lane & 31;
  for (int sg = blockIdx.x; sg < p.n_seg; sg += gridDim.x) {
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    float bp = 0.f;
    gram32<NO>(a, p.seg_beg[sg], p.seg_end[sg], half, comp, acc, bp);
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
  switch (a.rows.nother) {
    case 1: return launch_batched_t<1>(a, p, s);
    case 2: return launch_batched_t<2>(a, p, s);
    case 3: return launch_batched_t<3>(a, p, s);
    case 4: return launch_batched_t<4>(a, p, s);
    default: break;
  }
}

}  // namespace mgpu
