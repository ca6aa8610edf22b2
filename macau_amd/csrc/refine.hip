// refine.hip -- one step of iterative refinement for the batched sampler's
// ill-conditioned rows.
//
// The batched sampler factors P~ = Lambda + alpha G~ in FP32, with G~ the
// split-BF16 Gram, and returns u0 = L~^{-T}(L~^{-1} b + z), i.e. the solution
// of P~ u0 = b + L~ z.  For a row whose Cholesky pivots spread widely (the
// cheap condition estimate (max_i L_ii / min_i L_ii)^2 above a threshold) the
// rounding of G~ is amplified by the conditioning, so the epilogue queues
// the row with its factor L~, and k_refine applies one refinement step to
// the same system with the exact P = Lambda + alpha sum_e c_e c_e^T:
//   r = b + L~ z - P u0
//     = alpha sum_e c_e (y_e - mean - c_e . u0) + Lambda mu_n - Lambda u0 + L~ z,
//   u1 = u0 + L~^{-T} L~^{-1} r.
// The data term is accumulated from per-observation residuals (one pass over
// the row's nonzeros; c_e . u0 in FP64, the sums in FP32 per thread and the
// threads' sums combined in FP64),
// so P is never formed and the residual carries no cancellation between b
// and P u0.  The mean part of the sample (P^{-1} b) then carries the accuracy
// of the residual instead of the split-BF16 Gram's; the noise part
// (L^{-T} z) keeps the factor's.
//
// The train SSE fused into the last half-sweep keeps the unrefined sample of
// a refined row (the change is second order in the correction).
//
// The queue needs no separate reset launch: the last work-group of k_refine
// to finish (a ticket counter beside the queue count, taken by the
// work-groups that have rows to refine, or by work-group 0 alone when the
// queue is empty) folds the half-sweep's count into the running totals and
// clears the count for the next one.
#include "batched.h"

namespace mgpu {

constexpr int kRefineThreads = 256;

template <int KP>
struct RefineLds {
  float Ls[KP][KP + 1];
  float u0[KP], zs[KP], rs[KP];
  float part[kRefineThreads / 64][KP];
};

// Data term sum_e c_e (r_e - c_e . u0) over nonzeros [x0, x1) of the row, one
// nonzero per thread per step (FP32 per thread, c_e . u0 in FP64); returns
// the work-group's sum of component t in FP64 on threads t < KP.  u0 must be
// in LDS; ends with the work-group synchronised.
template <int KP, int NO>
__device__ double refine_data_term(const SampleArgs& a, int x0, int x1, RefineLds<KP>& s) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  float acc[KP];
#pragma unroll
  for (int k = 0; k < KP; k++) acc[k] = 0.f;
  for (int x = x0 + t; x < x1; x += kRefineThreads) {
    float c[KP];
    const float* f0 = a.others.f[0] + (size_t)a.rows.idx[(size_t)x * NO] * KP;
#pragma unroll
    for (int k = 0; k < KP; k += 4) {
      const float4 v = *reinterpret_cast<const float4*>(f0 + k);
      c[k] = v.x;
      c[k + 1] = v.y;
      c[k + 2] = v.z;
      c[k + 3] = v.w;
    }
#pragma unroll
    for (int o = 1; o < NO; o++) {
      const float* fo = a.others.f[o] + (size_t)a.rows.idx[(size_t)x * NO + o] * KP;
#pragma unroll
      for (int k = 0; k < KP; k += 4) {
        const float4 v = *reinterpret_cast<const float4*>(fo + k);
        c[k] *= v.x;
        c[k + 1] *= v.y;
        c[k + 2] *= v.z;
        c[k + 3] *= v.w;
      }
    }
    double sd = 0.0;
#pragma unroll
    for (int k = 0; k < KP; k++) sd = fma((double)c[k], (double)s.u0[k], sd);
    const float w = (float)((double)a.rows.r[x] - sd);
#pragma unroll
    for (int k = 0; k < KP; k++) acc[k] = fmaf(c[k], w, acc[k]);
  }
  // combine the threads' sums: a wave sum per component, then the waves
#pragma unroll
  for (int k = 0; k < KP; k++) {
    float v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) s.part[wave][k] = v;
  }
  __syncthreads();
  double R = 0.0;
  if (t < KP) {
#pragma unroll
    for (int w = 0; w < kRefineThreads / 64; w++) R += (double)s.part[w][t];
  }
  __syncthreads();
  return R;
}

// The refinement of queue entry e (row `row`) from the data term R (thread
// t < KP holds component t): r = alpha R + Lambda mu_n - Lambda u0 + L~ z,
// d = L~^{-T} L~^{-1} r on the first wave (lane = row of L~), u1 = u0 + d.
// u0 must be in LDS; loads L~ and z itself.
template <int KP>
__device__ void refine_finish(const SampleArgs& a, int e, int row, double R, double alpha, int it,
                              RefineLds<KP>& s) {
  const RefineQueue& q = a.refine;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const float* Lq = q.L + (size_t)e * KP * KP;
  for (int x = t; x < KP * KP; x += kRefineThreads) s.Ls[x / KP][x % KP] = Lq[x];
  if (t < KP)
    s.zs[t] = t < a.K ? normal_f32(a.seed, (uint32_t)row, (uint32_t)t, (uint32_t)it, make_tag(TAG_LATENT, a.mode))
                      : 0.f;
  __syncthreads();
  if (t < KP) {
    const double lmu = a.Lmu_rows ? (double)a.Lmu_rows[(size_t)row * KP + t] : (double)a.Lmu_f[t];
    double lu = 0.0, lz = 0.0;
    for (int j = 0; j < KP; j++) lu += (double)a.Lambda_f[t * KP + j] * (double)s.u0[j];
    for (int j = 0; j <= t; j++) lz += (double)s.Ls[t][j] * (double)s.zs[j];
    s.rs[t] = t < a.K ? (float)(alpha * R + lmu - lu + lz) : 0.f;
  }
  __syncthreads();
  if (wave == 0) {
    float y = lane < KP ? s.rs[lane] : 0.f;
    for (int j = 0; j < KP; j++) {
      const float yj = __shfl(lane == j ? y / s.Ls[j][j] : 0.f, j, 64);
      if (lane == j) y = yj;
      else if (lane > j && lane < KP) y -= s.Ls[lane][j] * yj;
    }
    for (int j = KP - 1; j >= 0; j--) {
      const float dj = __shfl(lane == j ? y / s.Ls[j][j] : 0.f, j, 64);
      if (lane == j) y = dj;
      else if (lane < j) y -= s.Ls[j][lane] * dj;
    }
    if (lane < a.K) a.U[(size_t)row * KP + lane] = s.u0[lane] + y;
  }
  __syncthreads();
}

// One work-group per queued row when many rows are queued.  When at most
// half the grid's work-groups would have a row, each row's nonzeros are
// split into C = grid / rows contiguous chunks, one work-group each: the
// chunks' data terms go to q.part (FP64), and the row's last chunk to finish
// (a per-row ticket) sums them in chunk order and finishes the row, so a
// queued row of 10^6 nonzeros is not left to one work-group.
template <int KP, int NO>
__global__ void __launch_bounds__(kRefineThreads) k_refine(SampleArgs a) {
  const RefineQueue& q = a.refine;
  __shared__ RefineLds<KP> s;
  __shared__ int last;
  const int t = threadIdx.x;
  const int it = *a.iter;
  const double alpha = *a.alpha;
  const int queued = *(volatile const int*)q.count;
  const int nq = queued < q.cap ? queued : q.cap;
  const int G = (int)gridDim.x;
  const int C = (nq > 0 && 2 * nq <= G) ? G / nq : 1;  // work-groups per row
  // work-groups past the queued rows (or chunks) leave at once and take no
  // ticket (with an empty queue only work-group 0 takes one and clears the
  // count)
  const int nact = nq < 1 ? 1 : (C > 1 ? nq * C : (nq < G ? nq : G));
  if ((int)blockIdx.x >= nact) return;
  // work items: rows e = blockIdx.x, + G, ... (C = 1), or chunk
  // blockIdx.x % C of row blockIdx.x / C (C > 1)
  const int first = C == 1 ? (int)blockIdx.x : (int)blockIdx.x / C;
  const int stride = C == 1 ? G : nq;
  for (int e = first; e < nq; e += stride) {
    const int row = q.row[e];
    if (t < KP) s.u0[t] = a.U[(size_t)row * KP + t];
    __syncthreads();
    const int beg = a.rows.row_ptr[row], len = a.rows.row_ptr[row + 1] - beg;
    const int c = C == 1 ? 0 : (int)blockIdx.x % C;
    const int x0 = beg + (int)((long long)len * c / C), x1 = beg + (int)((long long)len * (c + 1) / C);
    double R = refine_data_term<KP, NO>(a, x0, x1, s);
    if (C > 1) {
      if (t < KP) q.part[(size_t)blockIdx.x * KP + t] = R;
      __syncthreads();
      if (t == 0) {
        __threadfence();
        last = atomicAdd(q.row_ticket + e, 1) == C - 1;
      }
      __syncthreads();
      if (!last) continue;
      // every chunk of the row has stored its sum: add them in chunk order
      __threadfence();
      R = 0.0;
      if (t < KP)
        for (int cc = 0; cc < C; cc++) R += *(volatile const double*)(q.part + (size_t)(e * C + cc) * KP + t);
      if (t == 0) *(volatile int*)(q.row_ticket + e) = 0;
    }
    refine_finish<KP>(a, e, row, R, alpha, it, s);
  }
  // every work-group has read the count before taking its ticket; the last
  // one folds the count into the totals and clears the queue
  if (t == 0) {
    __threadfence();
    if (atomicAdd(q.count + 1, 1) == nact - 1) {
      q.totals[0] += queued;
      q.totals[1] += nq;
      *(volatile int*)q.count = 0;
      *(volatile int*)(q.count + 1) = 0;
      __threadfence();
    }
  }
}

template <int KP>
static void launch_refine_kp(const SampleArgs& a, hipStream_t s) {
  const dim3 g(kRefineGrid), b(kRefineThreads);
  switch (a.rows.nother) {
    case 1: hipLaunchKernelGGL((k_refine<KP, 1>), g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_refine<KP, 2>), g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_refine<KP, 3>), g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_refine<KP, 4>), g, b, 0, s, a); break;
    default: break;
  }
}

void launch_refine(const SampleArgs& a, hipStream_t s) {
  if (!a.refine.count) return;
  if (a.KP == 32) launch_refine_kp<32>(a, s);
  else if (a.KP == 64) launch_refine_kp<64>(a, s);
}

}  // namespace mgpu
