// batched.h -- helpers shared by the batched (split-BF16 MFMA) samplers.
#pragma once
#include <stdexcept>
#include <string>

#include "common.h"
#include "kernels.h"

namespace mgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// one lane's nonzero of a staged block: indices into the other modes and
// the residual
template <int NO>
struct BlkReg {
  int ix[NO];
  float r;
};

template <int NO>
struct Factors {
  __amdgpu_buffer_rsrc_t rs[NO];
};

// buffer descriptors of the other modes' factors (CS floats per column);
// word 3 = 0x00020000 selects 32-bit data, raw (unswizzled) addressing
template <int NO, int CS>
__device__ __forceinline__ Factors<NO> make_factors(const FactorSet& f) {
  Factors<NO> r;
#pragma unroll
  for (int t = 0; t < NO; t++)
    r.rs[t] = __builtin_amdgcn_make_buffer_rsrc((void*)f.f[t], (short)0, f.n[t] * CS * (int)sizeof(float),
                                                0x00020000);
  return r;
}

// component comp of factor column ix; ix = -1 gives an offset past
// num_records and the buffer load returns 0
template <int CS>
__device__ __forceinline__ float ld_factor(__amdgpu_buffer_rsrc_t rs, int ix, int comp) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, ix * (CS * 4) + comp * 4, 0, 0));
}

// x = hi + lo with hi = bf16(x) (round to nearest) and lo = bf16(x - hi),
// eight values packed in the MFMA operand order (element 2p in the low half
// of register p).  hi is converted once, packed (v_cvt_pk_bf16_f32), and
// widened back by a shift or a mask of the packed word; the empty asm keeps
// the compiler from re-deriving the low half's widening as a second
// conversion.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  const bf16x2 v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ void split8(const float* x, bf16x8& hi, bf16x8& lo) {
  uint32_t h[4], l[4];
#pragma unroll
  for (int p = 0; p < 4; p++) {
    h[p] = pack_bf16(x[2 * p], x[2 * p + 1]);
    asm volatile("" : "+v"(h[p]));
    const float h0 = __uint_as_float(h[p] << 16), h1 = __uint_as_float(h[p] & 0xffff0000u);
    l[p] = pack_bf16(x[2 * p] - h0, x[2 * p + 1] - h1);
  }
  hi = __builtin_bit_cast(bf16x8, h);
  lo = __builtin_bit_cast(bf16x8, l);
}

// nonzero e0 + lane of [e0, end) for the first BLK lanes (index -1 and
// residual 0 past the end)
template <int NO, int BLK>
__device__ __forceinline__ void load_blk(const SampleArgs& a, int e0, int end, int lane, BlkReg<NO>& b) {
  const int e = e0 + lane;
  if ((BLK == 64 || lane < BLK) && e < end) {
#pragma unroll
    for (int t = 0; t < NO; t++) b.ix[t] = a.rows.idx[(size_t)e * NO + t];
    b.r = a.rows.r[e];
  } else {
#pragma unroll
    for (int t = 0; t < NO; t++) b.ix[t] = -1;
    b.r = 0.f;
  }
}

// Sums each heavy slot's segment partials (W floats per segment, stored W
// apart) in segment order into the slot's first segment.  Thread (slot, q)
// owns entry q of the slot, so each segment's loads are coalesced across the
// work-group and the slots' sums run in parallel.
template <int W>
__global__ void __launch_bounds__(256) k_seg_sum(float* partial, const int* seg0, int nslot) {
  constexpr int CH = (W + 255) / 256;  // work-groups per slot
  const int slot = blockIdx.x / CH, q = (blockIdx.x % CH) * 256 + threadIdx.x;
  if (slot >= nslot || q >= W) return;
  const int s0 = seg0[slot], s1 = seg0[slot + 1];
  float* base = partial + (size_t)s0 * W + q;
  float acc = base[0];
  int sg = s0 + 1;
  for (; sg + 8 <= s1; sg += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = partial[(size_t)(sg + u) * W + q];
#pragma unroll
    for (int u = 0; u < 8; u++) acc += v[u];
  }
  for (; sg < s1; sg++) acc += partial[(size_t)sg * W + q];
  base[0] = acc;
}

template <int W>
inline void launch_seg_sum(const BatchPlan& p, hipStream_t s) {
  constexpr int CH = (W + 255) / 256;
  if (p.n_heavy > 0)
    hipLaunchKernelGGL((k_seg_sum<W>), dim3(p.n_heavy * CH), dim3(256), 0, s, p.partial, p.heavy_seg0, p.n_heavy);
}

// fork of the heavy-row segment kernels onto SegStreams::stream and their join
inline void seg_check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("batched sampler: stream fork/join: ") + hipGetErrorString(e));
}
inline void seg_fork(hipStream_t s, const SegStreams& seg) {
  seg_check(hipEventRecord(seg.fork, s));
  seg_check(hipStreamWaitEvent(seg.stream, seg.fork, 0));
}
inline void seg_record_join(const SegStreams& seg) { seg_check(hipEventRecord(seg.join, seg.stream)); }
inline void seg_wait_join(hipStream_t s, const SegStreams& seg) { seg_check(hipStreamWaitEvent(s, seg.join, 0)); }

// Entry r (per lane, 0..15) of a 16-register accumulator fragment.
__device__ __forceinline__ float frag_at(const f32x16& acc, int r) {
  float v = acc[0];
#pragma unroll
  for (int k = 1; k < 16; k++) v = r == k ? acc[k] : v;
  return v;
}

// Diagonal entry (j, j), j = lane % 32, of a 32 x 32 block in the MFMA
// accumulator layout (lane l, register r holds entry (8(r/4) + 4(l/32) + r%4,
// l%32)): lane j + 32 ((j >> 2) & 1) holds it in register 4 (j >> 3) + (j & 3).
__device__ __forceinline__ float frag_diag(const f32x16& acc, int lane) {
  const int j = lane & 31;
  const float mine = frag_at(acc, 4 * (j >> 3) + (j & 3));  // the diagonal on the holding lanes
  return __shfl(mine, j + 32 * ((j >> 2) & 1), 64);
}

// Condition estimate of one row from three probe solves y = L^{-1} D^{1/2} x
// done beside the forward solve of the RHS (D = diag(P); x = the row's
// normals z, all ones, alternating signs): |y|^2 = x^T Pt^{-1} x for the
// Jacobi-scaled Pt = D^{-1/2} P D^{-1/2} (unit diagonal, largest eigenvalue
// between 1 and K), and est = K max_x |y|^2 / |x|^2.  It is K for a diagonal
// P and about kappa(Pt) when a direction is nearly singular (a probe catches
// about 1/K of a weak direction), with three probes so that a weak direction
// orthogonal to one probe is still seen.  The scaling keeps rows whose
// spectrum is merely spread (one large eigenvalue along a common mean
// direction of the factors) from reading as ill-conditioned.  LANES lanes
// hold the row (32: a half of the wave; 64: the whole wave); each passes y_i
// of the three probes and z_i (zero for i >= K).
template <int LANES>
__device__ __forceinline__ float probe_estimate(int K, float y1, float y2, float y3, float z) {
  float s1 = y1 * y1, s2 = y2 * y2, s3 = y3 * y3, zz = z * z;
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
    s3 += __shfl_xor(s3, off, 64);
    zz += __shfl_xor(zz, off, 64);
  }
  const float kk = (float)K;
  return kk * fmaxf(fmaxf(s1 / fmaxf(zz, 1e-30f), s2 / kk), s3 / kk);
}

// The row's entry in the refinement queue when its estimate exceeds the
// threshold (and it has at least min_nnz nonzeros).  LANES as above; lane i
// owns row i of L, whose entries 0..i are at Lrow (the triangular layout of
// the epilogues); est is the same on all the row's lanes.
template <int KP, int LANES>
__device__ __forceinline__ void refine_enqueue(const SampleArgs& a, int row, int nnz, int i, int base_lane,
                                               float est, const float* Lrow) {
  const RefineQueue& q = a.refine;
  if (!q.count) return;
  const bool flag = row >= 0 && nnz >= q.min_nnz && est > q.thresh;
  int slot = -1;
  if (flag && i == 0) {
    slot = atomicAdd(q.count, 1);
    if (slot < q.cap) q.row[slot] = row;
  }
  slot = __shfl(slot, base_lane, 64);
  if (flag && slot >= 0 && slot < q.cap) {
    float* dst = q.L + (size_t)slot * KP * KP + (size_t)i * KP;
#pragma unroll
    for (int k = 0; k < KP; k += 4) {
      float4 v;
      v.x = k <= i ? Lrow[k] : 0.f;
      v.y = k + 1 <= i ? Lrow[k + 1] : 0.f;
      v.z = k + 2 <= i ? Lrow[k + 2] : 0.f;
      v.w = k + 3 <= i ? Lrow[k + 3] : 0.f;
      *reinterpret_cast<float4*>(dst + k) = v;
    }
  }
}

// wave-sum of the SSE shares into sse_partial[slot]
__device__ inline void write_sse(const SampleArgs& a, double acc, int slot) {
  if (!a.sse_partial) return;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x == 0) a.sse_partial[slot] = acc;
}

}  // namespace mgpu
