// linop.hip -- side-information link matrix: SpMM with K right-hand sides and
// the block conjugate gradient of solve_blockcg (linop.h:228-343), FP64.
//
// Vectors of the CG are stored feature-major: column f holds the nr
// right-hand-side entries of feature f contiguously (the reference's
// row-major [nrhs x nfeat] Eigen matrices, linop.h:249-266).  The SpMM
// F P^T (A_mul_Bx<N>, linop.h:374-403) gathers one such column per nonzero
// with NP lanes, 64/NP nonzeros per wave-step: an HBM/L2 gather-bound loop.
#include <algorithm>
#include <stdexcept>

#include "common.h"
#include "kernels.h"

namespace mgpu {

constexpr int kSpmmThreads = 256;

// four doubles: an FP64 matrix-core accumulator, or a 32-byte vector access
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 ld4(const double* p) {
  const double2 a = reinterpret_cast<const double2*>(p)[0], b = reinterpret_cast<const double2*>(p)[1];
  return dbl4{a.x, a.y, b.x, b.y};
}

__device__ __forceinline__ void st4(double* p, dbl4 v) {
  reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}


// one wave per row (matrices without a work list)
template <int NP>
__global__ void __launch_bounds__(kSpmmThreads)
    k_spmm(CsrDev A, const double* __restrict__ in, int nr, double* __restrict__ out, const double* __restrict__ reg,
           const double* __restrict__ add, const double* skip) {
  if (skip && *skip != 0.0) return;
  constexpr int S = 64 / NP;
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * kSpmmThreads + threadIdx.x) >> 6;
  const int nw = (gridDim.x * kSpmmThreads) >> 6;
  const int sub = lane / NP;
  for (int r = w; r < A.nrow; r += nw) {
    const int beg = A.row_ptr[r], end = A.row_ptr[r + 1];
    for (int c0 = 0; c0 < nr; c0 += NP) {  // one pass when nr <= NP
      const int comp = c0 + lane % NP;
      double acc = 0.0;
      if (comp < nr) {
        for (int e = beg + sub; e < end; e += S) {
          const double v = A.vals ? A.vals[e] : 1.0;
          acc += v * in[(size_t)A.cols[e] * nr + comp];
        }
      }
#pragma unroll
      for (int off = NP; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
      if (sub == 0 && comp < nr) {
        if (add) acc += (*reg) * add[(size_t)r * nr + comp];
        out[(size_t)r * nr + comp] = acc;
      }
    }
  }
}

// one wave per segment of at most kSpmmSegNnz nonzeros: power-law row
// lengths (a popular feature appears in most samples) are split so no wave
// walks a long row alone
template <int NP>
__global__ void __launch_bounds__(kSpmmThreads)
    k_spmm_seg(CsrDev A, const double* __restrict__ in, int nr, double* __restrict__ out,
               const double* __restrict__ reg, const double* __restrict__ add, const double* skip) {
  if (skip && *skip != 0.0) return;
  constexpr int S = 64 / NP;
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * kSpmmThreads + threadIdx.x) >> 6;
  const int nw = (gridDim.x * kSpmmThreads) >> 6;
  const int sub = lane / NP;
  for (int sg = w; sg < A.nseg; sg += nw) {
    const int4 d = A.seg[sg];  // row, first nnz, end nnz, slot
    for (int c0 = 0; c0 < nr; c0 += NP) {  // one pass when nr <= NP
      const int comp = c0 + lane % NP;
      double acc = 0.0;
      if (comp < nr) {
#pragma unroll 4
        for (int e = d.y + sub; e < d.z; e += S) {
          const double v = A.vals ? A.vals[e] : 1.0;
          acc += v * in[(size_t)A.cols[e] * nr + comp];
        }
      }
#pragma unroll
      for (int off = NP; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
      if (sub == 0 && comp < nr) {
        if (d.w < 0) {
          if (add) acc += (*reg) * add[(size_t)d.x * nr + comp];
          out[(size_t)d.x * nr + comp] = acc;
        } else {
          A.partial[(size_t)d.w * nr + comp] = acc;
        }
      }
    }
  }
}

// The same for an even nr <= 32 with 16-byte gathers: lanes (grp, cp), grp
// = lane / 16 a lane group, cp = lane % 16 a pair of components.  The work
// list holds the split rows' segments first, then the whole rows by
// decreasing length, so its first A.nlong segments (those of >= 32
// nonzeros) form a prefix:
//   * a long segment takes a whole wave, four nonzeros per gather step (one
//     per group): each batch of 64 column indices is one coalesced load
//     handed to the gathering lanes by ds_bpermute, a lane's 16 gathers of
//     a batch are issued together, the next batch's indices load meanwhile,
//     and the groups' sums are combined by two butterfly steps;
//   * short segments (a transposed feature matrix: most features occur in a
//     handful of rows or none) go four to a wave, one per group, each group
//     walking its own segment 16 indices per load and 8 gathers at a time.
// Past a segment's end a lane gathers row 0 (valid) with weight 0, and lanes
// past nr read a valid pair and discard it, so no gather sits behind a
// branch.  The row's regularisation operand is loaded before the gathers;
// it and the output row are streamed with nontemporal accesses, which
// leaves more of the L2 to the gathered table (C3: +1%).
__device__ __forceinline__ void spmm_store(const CsrDev& A, int4 d, int nr, int cp, double ax, double ay, double rg,
                                           double2 ad, double* __restrict__ out) {
  if (d.w < 0) {
    double* o = out + (size_t)d.x * nr + 2 * cp;
    __builtin_nontemporal_store(ax + rg * ad.x, o);
    __builtin_nontemporal_store(ay + rg * ad.y, o + 1);
  }
  else
    *reinterpret_cast<double2*>(A.partial + (size_t)d.w * nr + 2 * cp) = make_double2(ax, ay);
}

__device__ __forceinline__ void spmm_long(const CsrDev& A, int sg, const double* __restrict__ in, int nr,
                                          double* __restrict__ out, double rg, const double* __restrict__ add,
                                          int lane) {
  const int grp = lane >> 4, cp = lane & 15;
  const bool live = 2 * cp < nr;
  const int cpo = live ? 2 * cp : 0;
  const int4 d = A.seg[sg];  // row, first nnz, end nnz, slot
  double2 ad = make_double2(0.0, 0.0);
  if (add && d.w < 0 && live) {
    const double* a = add + (size_t)d.x * nr + 2 * cp;
    ad = make_double2(__builtin_nontemporal_load(a), __builtin_nontemporal_load(a + 1));
  }
  int myc = d.y + lane < d.z ? A.cols[d.y + lane] : 0;
  double myv = d.y + lane < d.z ? (A.vals ? A.vals[d.y + lane] : 1.0) : 0.0;
  double ax = 0.0, ay = 0.0;
  for (int base = d.y; base < d.z; base += 64) {
    const int en = base + 64 + lane;
    const int nc = en < d.z ? A.cols[en] : 0;
    const double nv = en < d.z ? (A.vals ? A.vals[en] : 1.0) : 0.0;
    double2 xs[16];
    double vs[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = 4 * u + grp;
      const int c = __shfl(myc, e, 64);
      vs[u] = __shfl(myv, e, 64);
      xs[u] = *reinterpret_cast<const double2*>(in + (size_t)c * nr + cpo);
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
      ax += vs[u] * xs[u].x;
      ay += vs[u] * xs[u].y;
    }
    myc = nc;
    myv = nv;
  }
  ax += __shfl_xor(ax, 16, 64);
  ay += __shfl_xor(ay, 16, 64);
  ax += __shfl_xor(ax, 32, 64);
  ay += __shfl_xor(ay, 32, 64);
  if (grp == 0 && live) spmm_store(A, d, nr, cp, ax, ay, rg, ad, out);
}

__device__ __forceinline__ void spmm_short4(const CsrDev& A, int first, const double* __restrict__ in, int nr,
                                            double* __restrict__ out, double rg, const double* __restrict__ add,
                                            int lane, int end) {
  const int grp = lane >> 4, cp = lane & 15;
  const bool live = 2 * cp < nr;
  const int cpo = live ? 2 * cp : 0;
  const int sg = first + grp;
  const bool has = sg < end;
  const int4 d = has ? A.seg[sg] : make_int4(0, 0, 0, -1);
  double2 ad = make_double2(0.0, 0.0);
  if (add && has && d.w < 0 && live) {
    const double* a = add + (size_t)d.x * nr + 2 * cp;
    ad = make_double2(__builtin_nontemporal_load(a), __builtin_nontemporal_load(a + 1));
  }
  const int len = d.z - d.y;
  int mx = max(len, __shfl_xor(len, 16, 64));
  mx = max(mx, __shfl_xor(mx, 32, 64));  // the longest of the wave's four segments
  int myc = cp < len ? A.cols[d.y + cp] : 0;
  double myv = cp < len ? (A.vals ? A.vals[d.y + cp] : 1.0) : 0.0;
  double ax = 0.0, ay = 0.0;
  for (int b = 0; b < mx; b += 16) {
    const int en = b + 16 + cp;
    const int nc = en < len ? A.cols[d.y + en] : 0;
    const double nv = en < len ? (A.vals ? A.vals[d.y + en] : 1.0) : 0.0;
    const int tc = min(16, mx - b);
    for (int t0 = 0; t0 < tc; t0 += 8) {
      double2 xs[8];
      double vs[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int src = 16 * grp + t0 + u;  // t0 + u < 16
        const int c = __shfl(myc, src, 64);
        vs[u] = __shfl(myv, src, 64);
        xs[u] = *reinterpret_cast<const double2*>(in + (size_t)c * nr + cpo);
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        ax += vs[u] * xs[u].x;
        ay += vs[u] * xs[u].y;
      }
    }
    myc = nc;
    myv = nv;
  }
  if (has && live) spmm_store(A, d, nr, cp, ax, ay, rg, ad, out);
}

// empty rows (the tail of the work list): out = reg * add, sixteen rows to
// a wave, four per group, their loads issued together
__device__ __forceinline__ void spmm_empty16(const CsrDev& A, int first, int nr, double* __restrict__ out, double rg,
                                             const double* __restrict__ add, int lane) {
  const int grp = lane >> 4, cp = lane & 15;
  const bool live = 2 * cp < nr;
  int row[4];
  double2 ad[4];
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const int sg = first + 4 * h + grp;
    row[h] = sg < A.nseg ? A.seg[sg].x : -1;
  }
#pragma unroll
  for (int h = 0; h < 4; h++)
    ad[h] = (add && row[h] >= 0 && live) ? *reinterpret_cast<const double2*>(add + (size_t)row[h] * nr + 2 * cp)
                                         : make_double2(0.0, 0.0);
#pragma unroll
  for (int h = 0; h < 4; h++)
    if (row[h] >= 0 && live)
      *reinterpret_cast<double2*>(out + (size_t)row[h] * nr + 2 * cp) = make_double2(rg * ad[h].x, rg * ad[h].y);
}

// work units: the A.nlong long segments, then groups of four short ones,
// then groups of sixteen empty rows
__device__ __forceinline__ void hyb_units(const CsrDev& A, const double* __restrict__ in, int nr,
                                          double* __restrict__ out, const double* __restrict__ reg,
                                          const double* __restrict__ add, int block, int nblock) {
  const int lane = threadIdx.x & 63;
  const int w = (block * kSpmmThreads + threadIdx.x) >> 6;
  const int nw = (nblock * kSpmmThreads) >> 6;
  const double rg = add ? *reg : 0.0;
  const int e0 = A.nseg - A.nempty;  // first empty row's entry
  const int nshort = (e0 - A.nlong + 3) / 4;
  const int nunit = A.nlong + nshort + (A.nempty + 15) / 16;
  for (int u = w; u < nunit; u += nw) {
    if (u < A.nlong) spmm_long(A, u, in, nr, out, rg, add, lane);
    else if (u < A.nlong + nshort) spmm_short4(A, A.nlong + 4 * (u - A.nlong), in, nr, out, rg, add, lane, e0);
    else spmm_empty16(A, e0 + 16 * (u - A.nlong - nshort), nr, out, rg, add, lane);
  }
}

__global__ void __launch_bounds__(kSpmmThreads)
    k_spmm_hyb(CsrDev A, const double* __restrict__ in, int nr, double* __restrict__ out,
               const double* __restrict__ reg, const double* __restrict__ add, const double* skip) {
  if (skip && *skip != 0.0) return;
  hyb_units(A, in, nr, out, reg, add, blockIdx.x, gridDim.x);
}

// rows split into several segments: a work-group per row sums the row's
// partials over 8 strands of slots (slot s on strand s % 8), the strands
// combined in strand order
__global__ void __launch_bounds__(kSpmmThreads)
    k_spmm_fix(CsrDev A, int nr, double* __restrict__ out, const double* __restrict__ reg,
               const double* __restrict__ add, const double* skip) {
  if (skip && *skip != 0.0) return;
  constexpr int NS = kSpmmThreads / 32;  // 8 strands of 32 components (nr <= 32)
  __shared__ double red[NS][32];
  const int comp = threadIdx.x & 31, st = threadIdx.x >> 5;
  for (int m = blockIdx.x; m < A.nmulti; m += gridDim.x) {
    const int s0 = A.multi_slot0[m], s1 = A.multi_slot0[m + 1];
    const int row = A.multi_row[m];
    for (int c0 = 0; c0 < nr; c0 += 32) {
      const int c = c0 + comp;
      double acc = 0.0;
      if (c < nr)
        for (int sl = s0 + st; sl < s1; sl += NS) acc += A.partial[(size_t)sl * nr + c];
      red[st][comp] = acc;
      __syncthreads();
      if (st == 0 && c < nr) {
        double v = 0.0;
#pragma unroll
        for (int h = 0; h < NS; h++) v += red[h][comp];
        if (add) v += (*reg) * add[(size_t)row * nr + c];
        out[(size_t)row * nr + c] = v;
      }
      __syncthreads();
    }
  }
}


// F P for a binary F with a dense column block, nr = 32 (dense_role 1): the
// dense columns through FP64 matrix-core products, the other nonzeros by
// gathers.  A work-group (8 waves) takes 64 rows of F; wave (tile, half)
// owns rows 16 tile .. + 15 and components 16 half .. + 15 of them.  The
// work-group first copies the in-rows of the dense columns into LDS as
// [half][j][16] (a k-step's four rows of 16 doubles: two bank halves, each
// read by two rows, the minimum for 64 lanes x 8 bytes).  MFMA step s
// multiplies the 16 x 4 bit block of the tile's rows x dense columns
// 4 s .. 4 s + 3 (lane l: row l % 16, column 4 s + l / 16, a bit of the
// row's words) by the 4 x 16 block of in-rows (lane l: column 4 s + l / 16,
// component l % 16).  Output slot v of lane l is row 4 v + l / 16,
// component l % 16.  The first kDenseRem remaining nonzeros of each row
// come from a fixed-width table, so their gathers are two loads deep and
// overlap the copy; rows with more take further batches after the
// products.  All loads that need no earlier load are issued before the
// convergence flag is tested.  Sums run in the matrix core's order, then
// the remaining nonzeros in CSR order.
constexpr int kDenseFThreads = 512;  // 8 waves: 64 rows of F per work-group
constexpr int kDenseStage = kDenseMaxCols * 16 / kDenseFThreads;  // 16-byte copies per thread

__global__ void __launch_bounds__(kDenseFThreads)
    k_spmm_dense_f(CsrDev A, const double* __restrict__ in, double* __restrict__ out, const double* skip) {
  __shared__ double sP[2 * kDenseMaxCols * 16];
  const double sk = skip ? *skip : 0.0;
  const int ndp = A.ndp;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, kq = lane >> 4;
  const int half = wv & 1;
  const int r0 = blockIdx.x * (kDenseFThreads / 8) + (wv >> 1) * 16;
  const int ra = min(r0 + m, A.nfrow - 1);  // rows past the end read the last row and are not stored
  // loads at clamped addresses, unused values dropped afterwards: no branch
  // around a load, so they are all in flight together
  uint32_t wd[kDenseMaxCols / 32];
#pragma unroll
  for (int u = 0; u < kDenseMaxCols / 32; u++) wd[u] = A.row_bits[(size_t)ra * A.ndw + min(u, A.ndw - 1)];
  int col[kDenseStage];
#pragma unroll
  for (int u = 0; u < kDenseStage; u++)
    col[u] = A.dcol[min((int)threadIdx.x + kDenseFThreads * u, ndp * 16 - 1) >> 4];
  // the lane's output rows 4 v + l / 16
  int rb[4], re[4];
  int4 rc[4][kDenseRem / 4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = min(r0 + 4 * v + kq, A.nfrow - 1);
    rb[v] = A.rem_ptr[row];
    re[v] = A.rem_ptr[row + 1];
#pragma unroll
    for (int t = 0; t < kDenseRem / 4; t++) rc[v][t] = reinterpret_cast<const int4*>(A.rem_fix + (size_t)row * kDenseRem)[t];
  }
  asm volatile("" ::: "memory");  // the loads above stay ahead of the flag's test
  if (sk != 0.0) return;
  double2 cv[kDenseStage];
#pragma unroll
  for (int u = 0; u < kDenseStage; u++) {
    const int e = threadIdx.x + kDenseFThreads * u;
    cv[u] = reinterpret_cast<const double2*>(in + (size_t)col[u] * 32)[e & 15];
  }
  double rv[4][kDenseRem];
#pragma unroll
  for (int v = 0; v < 4; v++)
#pragma unroll
    for (int t = 0; t < kDenseRem / 4; t++) {
      rv[v][4 * t + 0] = in[(size_t)rc[v][t].x * 32 + 16 * half + m];
      rv[v][4 * t + 1] = in[(size_t)rc[v][t].y * 32 + 16 * half + m];
      rv[v][4 * t + 2] = in[(size_t)rc[v][t].z * 32 + 16 * half + m];
      rv[v][4 * t + 3] = in[(size_t)rc[v][t].w * 32 + 16 * half + m];
    }
  // the copies' loads are all issued before the first of their guarded stores
#pragma unroll
  for (int u = 0; u < kDenseStage; u++) asm volatile("" ::"v"(cv[u].x), "v"(cv[u].y));
#pragma unroll
  for (int u = 0; u < kDenseStage; u++) {
    const int e = threadIdx.x + kDenseFThreads * u, j = e >> 4, p = e & 15;
    if (e < ndp * 16) *reinterpret_cast<double2*>(sP + ((size_t)(p >> 3) * ndp + j) * 16 + 2 * (p & 7)) = cv[u];
  }
#pragma unroll
  for (int v = 0; v < 4; v++) re[v] = r0 + 4 * v + kq < A.nfrow ? re[v] : rb[v];
#pragma unroll
  for (int u = 0; u < kDenseMaxCols / 32; u++) wd[u] = u < A.ndw ? wd[u] : 0u;
  __syncthreads();
  const double* sB = sP + (size_t)half * ndp * 16 + m;
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  // 16 dense columns (4 steps) at a time: their LDS reads, then the products
#pragma unroll
  for (int g = 0; g < kDenseMaxCols / 16; g++) {
    if (16 * g >= ndp) break;
    double b[4];
#pragma unroll
    for (int st = 0; st < 4; st++) b[st] = sB[(16 * g + 4 * st + kq) * 16];
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const double a = (wd[g >> 1] >> (16 * (g & 1) + 4 * st + kq)) & 1u ? 1.0 : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[st], acc, 0, 0, 0);
    }
  }
  double x[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    x[v] = acc[v];
#pragma unroll
    for (int t = 0; t < kDenseRem; t++)
      if (rb[v] + t < re[v]) x[v] += rv[v][t];
  }
  // rows with more: batches of kDenseRem for every row of the wave at once
  for (int e0 = kDenseRem;; e0 += kDenseRem) {
    bool left = false;
#pragma unroll
    for (int v = 0; v < 4; v++) left |= rb[v] + e0 < re[v];
    if (!__any(left)) break;
    int c2[4][kDenseRem];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int t = 0; t < kDenseRem; t++) c2[v][t] = A.rem_cols[max(min(rb[v] + e0 + t, re[v] - 1), 0)];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int t = 0; t < kDenseRem; t++) rv[v][t] = in[(size_t)c2[v][t] * 32 + 16 * half + m];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int t = 0; t < kDenseRem; t++)
        if (rb[v] + e0 + t < re[v]) x[v] += rv[v][t];
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = r0 + 4 * v + kq;
    if (row < A.nfrow) out[(size_t)row * 32 + 16 * half + m] = x[v];
  }
}

// partial sums of F^T Y over the dense columns of a binary F, nr = 32
// (dense_role 2 on F^T).  Work-group (group g, half h, chunk c) copies the
// kDenseRowChunk rows of Y in chunk c, components 16 h .. 16 h + 15, into
// LDS ([row][16]: a step's four rows fall two to a bank half), and its wave
// w multiplies the 16 x kDenseRowChunk bit block of dense columns
// 16 (4 g + w) .. + 15 x those rows of F (lane l: column l % 16, row
// 4 s + l / 16 of step s, from the column's bit words) by them (lane l: row
// 4 s + l / 16, component l % 16), writing slot dslot0 + j nkc + c of each
// dense column j; k_spmm_fix adds the chunks in order (and reg * add).  Rows
// past the end have no bits; their copies read the last row.
__device__ __forceinline__ void dense_ft_block(const CsrDev& A, const double* __restrict__ in, int bid,
                                               double* sY) {
  constexpr int KS = kDenseRowChunk / 4, NW = kDenseRowChunk / 32;
  constexpr int CP = kDenseRowChunk * 8 / 256;  // 16-byte copies per thread
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, kq = lane >> 4;
  const int kc = bid % A.nkc, half = (bid / A.nkc) & 1, g = bid / A.nkc >> 1;
  const int d0 = (4 * g + wv) * 16;
  const bool active = d0 < A.ndp;  // a uniform per wave
  const int rb = kc * kDenseRowChunk;
  double2 cv[CP];
#pragma unroll
  for (int u = 0; u < CP; u++) {
    const int e = threadIdx.x + 256 * u, row = min(rb + (e >> 3), A.nfrow - 1);
    cv[u] = reinterpret_cast<const double2*>(in + (size_t)row * 32 + 16 * half)[e & 7];
  }
  uint32_t wd[NW];
#pragma unroll
  for (int u = 0; u < NW; u++)
    wd[u] = A.col_bits[(size_t)min(d0 + m, A.ndp - 1) * A.nrw + min(rb / 32 + u, A.nrw - 1)];
#pragma unroll
  for (int u = 0; u < CP; u++) {
    const int e = threadIdx.x + 256 * u;
    *reinterpret_cast<double2*>(sY + (e >> 3) * 16 + 2 * (e & 7)) = cv[u];
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int u = 0; u < NW; u++) wd[u] = rb / 32 + u < A.nrw ? wd[u] : 0u;
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int gs = 0; gs < KS / 8; gs++) {
    double b[8];
#pragma unroll
    for (int st = 0; st < 8; st++) b[st] = sY[(32 * gs + 4 * st + kq) * 16 + m];
#pragma unroll
    for (int st = 0; st < 8; st++) {
      const double a = (wd[gs] >> (4 * st + kq)) & 1u ? 1.0 : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[st], acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int j = d0 + 4 * v + kq;
    if (j < A.nd) A.partial[((size_t)A.dslot0 + (size_t)j * A.nkc + kc) * 32 + 16 * half + m] = acc[v];
  }
}

// F^T Y + reg add for the dense-block plan of F^T in one launch: the first
// ndense work-groups take the dense columns' chunk sums (above), the rest
// the other rows' work units (k_spmm_hyb); k_spmm_fix then completes the
// split rows and the dense columns
__global__ void __launch_bounds__(256)
    k_spmm_dense_ft(CsrDev A, const double* __restrict__ in, int nr, double* __restrict__ out,
                    const double* __restrict__ reg, const double* __restrict__ add, int ndense, const double* skip) {
  __shared__ double sY[kDenseRowChunk * 16];
  if (skip && *skip != 0.0) return;
  if ((int)blockIdx.x < ndense) dense_ft_block(A, in, blockIdx.x, sY);
  else hyb_units(A, in, nr, out, reg, add, blockIdx.x - ndense, gridDim.x - ndense);
}

template <int NP>
static void launch_spmm_t(const CsrDev& A, const double* in, int nr, double* out, const double* reg,
                          const double* add, hipStream_t s, const double* skip) {
  const bool aligned = (((uintptr_t)in | (uintptr_t)out | (uintptr_t)add | (uintptr_t)A.partial) & 15) == 0;
  if (A.dense_role == 1 && nr == 32 && !add && aligned) {
    hipLaunchKernelGGL(k_spmm_dense_f, dim3((A.nfrow + kDenseFThreads / 8 - 1) / (kDenseFThreads / 8)), dim3(kDenseFThreads), 0,
                       s, A, in, out, skip);
    return;
  }
  if (A.dense_role == 2 && nr == 32 && aligned && A.nd > 0) {
    const int ndense = (A.ndp / 16 + 3) / 4 * 2 * A.nkc;  // four column tiles per work-group
    const int nunit = A.nlong + (A.nseg - A.nempty - A.nlong + 3) / 4 + (A.nempty + 15) / 16;
    const int g = A.nseg > 0 ? std::min((nunit + 3) / 4, 16384) : 0;
    hipLaunchKernelGGL(k_spmm_dense_ft, dim3(ndense + g), dim3(256), 0, s, A, in, nr, out, reg, add, ndense, skip);
    if (A.nmulti > 0) {
      const int g2 = A.nmulti < 4096 ? A.nmulti : 4096;
      hipLaunchKernelGGL(k_spmm_fix, dim3(g2), dim3(kSpmmThreads), 0, s, A, nr, out, reg, add, skip);
    }
    return;
  }
  if (A.dense_role != 0) throw std::logic_error("dense-block SpMM plan used with nr != 32 or unaligned vectors");
  if (A.nseg > 0) {
    int grid = (A.nseg + 3) / 4;
    if (grid > 16384) grid = 16384;
    if (NP == 32 && nr % 2 == 0 && aligned) {
      const int nunit = A.nlong + (A.nseg - A.nempty - A.nlong + 3) / 4 + (A.nempty + 15) / 16;
      int g = (nunit + 3) / 4;
      if (g > 16384) g = 16384;
      hipLaunchKernelGGL(k_spmm_hyb, dim3(g), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add, skip);
    } else
      hipLaunchKernelGGL((k_spmm_seg<NP>), dim3(grid), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add, skip);
    if (A.nmulti > 0) {
      const int g2 = A.nmulti < 4096 ? A.nmulti : 4096;
      hipLaunchKernelGGL(k_spmm_fix, dim3(g2), dim3(kSpmmThreads), 0, s, A, nr, out, reg, add, skip);
    }
    return;
  }
  int grid = (A.nrow + 3) / 4;
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL((k_spmm<NP>), dim3(grid), dim3(kSpmmThreads), 0, s, A, in, nr, out, reg, add, skip);
}

void launch_spmm(const CsrDev& A, const double* in, int nr, double* out, const double* reg, const double* add,
                 hipStream_t s, const double* skip) {
  if (A.nrow == 0) return;
  if (nr <= 8) launch_spmm_t<8>(A, in, nr, out, reg, add, s, skip);
  else if (nr <= 16) launch_spmm_t<16>(A, in, nr, out, reg, add, s, skip);
  else if (nr <= 32) launch_spmm_t<32>(A, in, nr, out, reg, add, s, skip);
  else launch_spmm_t<64>(A, in, nr, out, reg, add, s, skip);
}

// ------------------------------------------------------------------
// dot products  sum_f a[:,f] b[:,f]^T  (A_mul_Bt_omp_sym / A_mul_At_omp,
// linop.cpp:271-372), FP64.  Same structure as the hyperprior statistics:
// column groups stream double-buffered LDS tiles of 8 columns of a and b,
// thread (ti, tj) of a group owns a 4 x 4 tile of the nr x nr sum; the
// groups of a block are combined in order, then the block partials are
// reduced in place (launch_partial_reduce).
// ------------------------------------------------------------------
constexpr int kMaxNr = 40;

int dots_blocks(int n) {
  int b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 512) b = 512;
  return b;
}

template <int NRP>
struct DotShape {
  static constexpr int TD = NRP / 4;
  static constexpr int GT = TD * TD;
  static constexpr int G = GT >= 128 ? 1 : ((256 / GT) < 4 ? (256 / GT) : 4);
  static constexpr int THREADS = G * GT;
  static constexpr int TC = 8;
  static constexpr int TILE = TC * NRP;
  static constexpr int EPT = (TILE + GT - 1) / GT;
  static constexpr int RED = G > 1 ? G * NRP * NRP : 0;
  static constexpr int SMEM = (G * 4 * TILE > RED) ? G * 4 * TILE : RED;
};

template <int NRP>
__global__ void __launch_bounds__(DotShape<NRP>::THREADS)
    k_dots(const double* __restrict__ a, const double* __restrict__ b, int nr, int n, int cols,
           double* __restrict__ partial, const double* skip) {
  if (skip && *skip != 0.0) return;
  using S = DotShape<NRP>;
  constexpr int TD = S::TD, GT = S::GT, G = S::G, TC = S::TC, TILE = S::TILE, EPT = S::EPT;
  __shared__ double smem[S::SMEM];
  const int t = threadIdx.x, g = t / GT, gt = t % GT;
  const int ti = gt % TD, tj = gt / TD;
  const int b0 = blockIdx.x * cols, b1 = min(n, b0 + cols);
  const int per = (cols + G - 1) / G;
  const int c0 = min(b1, b0 + g * per), c1 = min(b1, c0 + per);
  const int ntiles = (per + TC - 1) / TC;
  double* ta = smem + (size_t)g * 4 * TILE;  // [2][TILE] of a, then [2][TILE] of b
  double ra[EPT], rb[EPT];
  auto load = [&](int tt) {
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int f = gt + e * GT, c = f / NRP, i = f % NRP;
      const int col = c0 + tt * TC + c;
      const bool ok = f < TILE && col < c1 && i < nr;
      ra[e] = ok ? a[(size_t)col * nr + i] : 0.0;
      rb[e] = ok ? b[(size_t)col * nr + i] : 0.0;
    }
  };
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[r][q] = 0.0;
  load(0);
  for (int tt = 0; tt < ntiles; tt++) {
    double* xa = ta + (tt & 1) * TILE;
    double* xb = ta + 2 * TILE + (tt & 1) * TILE;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const int f = gt + e * GT;
      if (f < TILE) {
        xa[f] = ra[e];
        xb[f] = rb[e];
      }
    }
    __syncthreads();
    if (tt + 1 < ntiles) load(tt + 1);
#pragma unroll 2
    for (int c = 0; c < TC; c++) {
      double ai[4], bj[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        ai[r] = xa[c * NRP + 4 * ti + r];
        bj[r] = xb[c * NRP + 4 * tj + r];
      }
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[r][q] += ai[r] * bj[q];
    }
  }
  const int NN = nr * nr;
  double* out = partial + (size_t)blockIdx.x * NN;
  if (G == 1) {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = 4 * ti + r, j = 4 * tj + q;
        if (i < nr && j < nr) out[i + j * nr] = acc[r][q];
      }
    return;
  }
  __syncthreads();  // tiles are dead: reuse the LDS for the group sums
  double* red = smem + (size_t)g * NN;
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = 4 * ti + r, j = 4 * tj + q;
      if (i < nr && j < nr) red[i + j * nr] = acc[r][q];
    }
  __syncthreads();
  for (int q = t; q < NN; q += S::THREADS) {
    double v = 0.0;
#pragma unroll
    for (int h = 0; h < G; h++) v += smem[(size_t)h * NN + q];
    out[q] = v;
  }
}

// The same sums at nr = 32 on the FP64 matrix cores: v_mfma_f64_16x16x4f64
// with the features as the reduction index, four per instruction.  Lane
// (q = l / 16, j = l % 16) loads components 2j and 2j + 1 of feature f + q
// of a and of b as one 16-byte vector; the even components feed the tiles
// of row (column) block 0 and the odd ones block 1, so the four tiles
// (ib, jb) hold D[2i + ib][2j' + jb] and every load is a 256-byte row read
// by 16 lanes.  Each wave takes a quarter of the block's columns; the four
// waves' sums are added in wave order through LDS.
__global__ void __launch_bounds__(256)
    k_dots32(const double* __restrict__ a, const double* __restrict__ b, int n, int cols,
             double* __restrict__ partial, const double* skip, const int* __restrict__ nxt,
             const double* __restrict__ reg) {
  if (skip && *skip != 0.0) return;
  __shared__ double red[4][32 * 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
  const int b0 = blockIdx.x * cols, b1 = min(n, b0 + cols);
  const int per = (cols + 3) / 4;
  const int c0 = min(b1, b0 + wv * per), c1 = min(b1, c0 + per);
  typedef double dbl4v __attribute__((ext_vector_type(4)));
  dbl4v d[2][2];
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++) d[u][v] = dbl4v{0.0, 0.0, 0.0, 0.0};
  // 32 features per trip: the 16 loads of a lane are issued together
  // (clamped addresses, zeros selected past the end: no branch).  With nxt,
  // a trip whose features have no nonzero in F^T takes b = reg * a (b is
  // K P there, not written), and nxt for the next trip loads ahead.
  const double lam = nxt ? *reg : 0.0;
  int e = nxt ? nxt[min(c0, n)] : 0;
  for (int f0 = c0; f0 < c1; f0 += 32) {
    const bool emp = nxt && e >= min(f0 + 32, c1);  // the same on every lane
    if (nxt) e = nxt[min(f0 + 32, n)];
    double2 xs[8], ys[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int fc = min(f0 + 4 * u + q, c1 - 1);
      xs[u] = reinterpret_cast<const double2*>(a + (size_t)fc * 32)[j];
    }
    if (emp) {
#pragma unroll
      for (int u = 0; u < 8; u++) ys[u] = make_double2(lam * xs[u].x, lam * xs[u].y);
    } else {
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int fc = min(f0 + 4 * u + q, c1 - 1);
        ys[u] = reinterpret_cast<const double2*>(b + (size_t)fc * 32)[j];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const bool ok = f0 + 4 * u + q < c1;
      const double x0 = ok ? xs[u].x : 0.0, x1 = ok ? xs[u].y : 0.0;
      const double y0 = ok ? ys[u].x : 0.0, y1 = ok ? ys[u].y : 0.0;
      d[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y0, d[0][0], 0, 0, 0);
      d[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y1, d[0][1], 0, 0, 0);
      d[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y0, d[1][0], 0, 0, 0);
      d[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y1, d[1][1], 0, 0, 0);
    }
  }
  // slot v of lane (q, j) in tile (ib, jb) is D[2 (4v + q) + ib][2j + jb]
#pragma unroll
  for (int ib = 0; ib < 2; ib++)
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[wv][(2 * (4 * v + q) + ib) + 32 * (2 * j + jb)] = d[ib][jb][v];
  __syncthreads();
  double* out = partial + (size_t)blockIdx.x * 1024;
  for (int e = threadIdx.x; e < 1024; e += 256) out[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

void launch_dots(const double* a, const double* b, int nr, int n, double* partial, int nblk, hipStream_t s,
                 const double* skip, bool reduce, const int* nxt, const double* reg) {
  const int cols = (n + nblk - 1) / nblk;
  if (nr == 32 && (((uintptr_t)a | (uintptr_t)b) & 15) == 0) {
    hipLaunchKernelGGL(k_dots32, dim3(nblk), dim3(256), 0, s, a, b, n, cols, partial, skip, nxt, reg);
    if (reduce && nblk > 1) launch_partial_reduce(partial, nblk, nr * nr, s);
    return;
  }
  if (nxt) throw std::logic_error("launch_dots: the reg * a shortcut needs nr = 32 and aligned vectors");
#define MGPU_DOTS(P)                                                                                        \
  hipLaunchKernelGGL((k_dots<P>), dim3(nblk), dim3(DotShape<P>::THREADS), 0, s, a, b, nr, n, cols, partial, \
                     skip)
  if (nr <= 8) MGPU_DOTS(8);
  else if (nr <= 16) MGPU_DOTS(16);
  else if (nr <= 32) MGPU_DOTS(32);
  else MGPU_DOTS(40);
#undef MGPU_DOTS
  if (reduce && nblk > 1) launch_partial_reduce(partial, nblk, nr * nr, s);
}

// ------------------------------------------------------------------
// small nr x nr steps (one wave)
// small layout: RtR | RtR2 | PtKP | A | Psi (nr*nr each) | norms (nr) |
//               flags: done, iters (as doubles)
// ------------------------------------------------------------------
struct SmallView {
  double *RtR, *RtR2, *PtKP, *A, *Psi, *norms, *flags;
  __device__ SmallView(double* s, int nr) {
    const int NN = nr * nr;
    RtR = s;
    RtR2 = s + NN;
    PtKP = s + 2 * NN;
    A = s + 3 * NN;
    Psi = s + 4 * NN;
    norms = s + 5 * NN;
    flags = norms + nr;
  }
};

// The small steps run on one wave: the LDS accesses of a wave are performed
// in order, so a wait for them and a compiler fence order its lanes' LDS
// stores before the following loads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  lds_wait();
}

// reduce partials -> lower-symmetric matrix in LDS
// The small-step kernels run one wave (64 threads); their nr x nr operands
// are moved with all loads of a thread issued together.
constexpr int kSmallPer = (kMaxNr * kMaxNr + 63) / 64;
// LDS workspace of a small solve: the LLT path's padded factor, or the
// elimination's rows of U (32 x 64) and reciprocal pivots (32)
constexpr int kSolveWork = 32 * 64 + 32;
static_assert(kMaxNr * (kMaxNr + 1) <= kSolveWork, "workspace holds the padded factor");

__device__ __forceinline__ void gather_nn(const double* __restrict__ src, double (&v)[kSmallPer], int NN) {
#pragma unroll
  for (int u = 0; u < kSmallPer; u++) {
    const int q = threadIdx.x + 64 * u;
    v[u] = q < NN ? src[q] : 0.0;
  }
}

__device__ __forceinline__ void scatter_nn(const double (&v)[kSmallPer], double* dst, int NN) {
#pragma unroll
  for (int u = 0; u < kSmallPer; u++) {
    const int q = threadIdx.x + 64 * u;
    if (q < NN) dst[q] = v[u];
  }
}

__device__ __forceinline__ void load_nn(const double* __restrict__ src, double* dst, int NN) {
  double v[kSmallPer];
  gather_nn(src, v, NN);
  scatter_nn(v, dst, NN);
}

// M = the symmetric matrix whose lower triangle is partial[0] (makeSymmetric,
// linop.cpp:241-243; the dots leave their total in the first partial)
__device__ __forceinline__ void reduce_sym(const double* partial, int nblk, int nr, double* M) {
  (void)nblk;
  const int NN = nr * nr;
  double v[kSmallPer];
  // (i, j) of entry q = lane + 64 u, stepped without divisions
  const int di = 64 % nr, dj = 64 / nr;
  int i = threadIdx.x % nr, j = threadIdx.x / nr;
#pragma unroll
  for (int u = 0; u < kSmallPer; u++) {
    const int q = threadIdx.x + 64 * u;
    v[u] = q < NN ? partial[(i >= j) ? q : (j + i * nr)] : 0.0;
    i += di;
    j += dj;
    if (i >= nr) {
      i -= nr;
      j++;
    }
  }
#pragma unroll
  for (int u = 0; u < kSmallPer; u++) {
    const int q = threadIdx.x + 64 * u;
    if (q < NN) M[q] = v[u];
  }
  wave_sync();
}

// X = (M^{-1} Q)^T with M factored as Eigen's unblocked LLT does it
// (left-looking; stops at the first non-positive pivot and solves with the
// partially factored lower triangle, whose unfactored columns keep M's
// values).  The reference's block CG relies on that behaviour once the
// block Krylov space is exhausted.  One wave, problem padded to NRP with an
// identity block:
//   * factorisation (left-looking, in LDS): step k forms
//     s_i = A[i][k] - sum_{j<k} L[i][j] L[k][j] on lane i, the pivot s_k
//     comes from lane k, and lanes i >= k write column k;
//   * solves: lane j holds column j of the right-hand side in registers and
//     runs the forward and back substitution with rows/columns of L read as
//     LDS broadcasts.
// M, Q are nr x nr col-major in LDS, X is written col-major (X[j][i] =
// S[i][j]); Ls is NRP x (NRP + 1) scratch (row-major).
template <int NRP>
__device__ __forceinline__ void small_solve_w(const double* M, const double* Q, double* X, int nr, double* Ls) {
  constexpr int LD = NRP + 1;  // padded row stride: lanes reading their own rows hit distinct banks
  int lane = threadIdx.x & 63;
  opaque(lane);
  const int i = lane < NRP ? lane : NRP - 1;  // lanes >= NRP shadow the last row
#pragma unroll
  for (int j = 0; j < NRP; j++) {
    const bool in = i < nr && j < nr;
    Ls[i * LD + j] = in ? M[i + j * nr] : (i == j ? 1.0 : 0.0);
  }
  lds_wait();
  for (int k = 0; k < NRP; k++) {
    double s = Ls[i * LD + k];
#pragma unroll 8
    for (int j = 0; j < k; j++) s -= Ls[i * LD + j] * Ls[k * LD + j];
    const double d = __shfl(s, k, 64);
    if (!(d > 0.0)) break;  // the pivot is the same on every lane
    const double lkk = sqrt(d);
    if (i == k) Ls[i * LD + k] = lkk;
    else if (i > k) Ls[i * LD + k] = s / lkk;
    lds_wait();
  }
  // forward L y = q and back L^T x = y, column j of the right-hand side on lane j
  const int j = lane < NRP ? lane : NRP - 1;
  double y[NRP];
#pragma unroll
  for (int m = 0; m < NRP; m++) y[m] = (j < nr && m < nr) ? Q[m + j * nr] : 0.0;
#pragma unroll
  for (int r = 0; r < NRP; r++) {
    int o = 0;  // opaque zero offset: the row's reads wait for the previous row
    opaque_after(o, r > 0 ? y[r - 1] : 0.0);
    double acc = y[r];
#pragma unroll
    for (int m = 0; m < r; m++) acc -= Ls[r * LD + m + o] * y[m];
    y[r] = acc / Ls[r * LD + r + o];
  }
#pragma unroll
  for (int r = NRP - 1; r >= 0; r--) {
    int o = 0;
    opaque_after(o, r < NRP - 1 ? y[r + 1] : 0.0);
    double acc = y[r];
#pragma unroll
    for (int m = r + 1; m < NRP; m++) acc -= Ls[m * LD + r + o] * y[m];
    y[r] = acc / Ls[r * LD + r + o];
  }
  if (lane < nr) {
#pragma unroll
    for (int r = 0; r < NRP; r++)
      if (r < nr) X[lane + r * nr] = y[r];
  }
  lds_wait();
}

// The same X = (M^{-1} Q)^T for nr <= 32 by Gaussian elimination of the
// augmented [M | Q] without pivoting: lane j holds column j of M (j < NRP,
// identity padding), lane 32 + j column j of Q, as NRP doubles in registers.
// Step k stores row k of the reduced matrix in LDS (row k of U in M = L'U);
// its entries right of the pivot are also the multipliers' numerators,
// since the trailing block of the reduced matrix stays symmetric (only the
// upper triangle is ever read).  Each step is one LDS round trip plus
// NRP - k - 1 independent FMAs per lane, so the 32 steps cost a few
// thousand cycles instead of the LLT path's chains of dependent LDS reads:
// row k + 1 is updated first and its LDS round trip overlaps the rest of
// step k, and the pivot's reciprocal is the hardware estimate refined by
// two Newton steps.  The back substitution reads U's columns above the
// pivot, which are all in LDS once the forward pass ends.  A pivot below
// 1e-10 of its diagonal entry of M (or NaN) returns false before anything
// is written: near a non-positive pivot the caller runs the LLT path, which
// reproduces Eigen's stop at a non-positive pivot.  U holds NRP x 64 + NRP
// doubles (the rows, then the pivot thresholds).
template <int NRP>
__device__ __forceinline__ bool small_solve_elim(const double* M, const double* Q, double* X, int nr, double* U) {
  static_assert(NRP <= 32, "two columns per lane pair");
  const int lane = threadIdx.x & 63, col = lane & 31;
  const bool isq = lane >= 32;
  double a[NRP];
#pragma unroll
  for (int r = 0; r < NRP; r++) {
    double v = 0.0;
    if (col < nr && r < nr) v = isq ? Q[r + col * nr] : M[r + col * nr];
    else if (!isq && r == col) v = 1.0;
    a[r] = v;
  }
  double* thr = U + NRP * 64;  // pivot thresholds 1e-10 M_kk
  if (lane < NRP) thr[lane] = lane < nr ? 1e-10 * M[lane + lane * nr] : 1e-10;
  U[lane] = a[0];
  lds_wait();
  double u[NRP], rp[NRP];
#pragma unroll
  for (int i = 0; i < NRP; i++) u[i] = U[i];
  double th = thr[0];
  bool bad = false;
#pragma unroll
  for (int k = 0; k < NRP; k++) {
    // u holds row k of the reduced matrix
    bad |= !(u[k] > th);  // checked once after the loop: no branch per step
    double inv = __builtin_amdgcn_rcp(u[k]);  // two Newton steps from the hardware estimate
    inv = fma(inv, fma(-u[k], inv, 1.0), inv);
    inv = fma(inv, fma(-u[k], inv, 1.0), inv);
    rp[k] = inv;
    const double t = a[k] * inv;
    double un[NRP], thn = 0.0;
    if (k + 1 < NRP) {
      a[k + 1] -= u[k + 1] * t;
      U[(k + 1) * 64 + lane] = a[k + 1];
      // a wave's LDS operations are performed in order, so the next row's
      // reads may follow its write at once; the fence only stops the
      // compiler from moving them
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
      for (int i = k + 1; i < NRP; i++) un[i] = U[(k + 1) * 64 + i];
      thn = thr[k + 1];
    }
#pragma unroll
    for (int i = k + 2; i < NRP; i++) a[i] -= u[i] * t;
    // apply this step's updates here rather than sink them to each a[i]'s
    // next use (which would keep every step's row alive in registers)
#pragma unroll
    for (int i = k + 2; i < NRP; i++) asm volatile("" : "+v"(a[i]));
#pragma unroll
    for (int i = k + 1; i < NRP; i++) u[i] = un[i];
    th = thn;
  }
  if (bad) return false;  // uniform: every lane read the same pivots
#pragma unroll
  for (int k = NRP - 1; k >= 0; k--) {
    int o = 0;  // opaque zero offset: step k's loads wait for step k + 2 (two steps of lookahead)
    opaque_after(o, k + 2 < NRP ? a[k + 2] : 0.0);
    a[k] *= rp[k];
#pragma unroll
    for (int i = 0; i < k; i++) a[i] -= U[i * 64 + k + o] * a[k];
  }
  if (isq && col < nr) {
#pragma unroll
    for (int r = 0; r < NRP; r++)
      if (r < nr) X[col + r * nr] = a[r];
  }
  lds_wait();
  return true;
}

__device__ __forceinline__ void small_solve(const double* M, const double* Q, double* X, int nr, double* Ls) {
  if (nr <= 8 && small_solve_elim<8>(M, Q, X, nr, Ls)) return;
  if (nr > 8 && nr <= 16 && small_solve_elim<16>(M, Q, X, nr, Ls)) return;
  if (nr > 16 && nr <= 32 && small_solve_elim<32>(M, Q, X, nr, Ls)) return;
  if (nr <= 8) small_solve_w<8>(M, Q, X, nr, Ls);
  else if (nr <= 16) small_solve_w<16>(M, Q, X, nr, Ls);
  else if (nr <= 32) small_solve_w<32>(M, Q, X, nr, Ls);
  else small_solve_w<40>(M, Q, X, nr, Ls);
}

// norms from diag(B B^T): small.norms; flags reset
__global__ void k_cg_norms(const double* partial, int nblk, int nr, double* small) {
  SmallView v(small, nr);
  const int NN = nr * nr;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[(size_t)b * NN + i + i * nr];
    v.norms[i] = sqrt(s);
  }
  if (threadIdx.x == 0) {
    v.flags[0] = 0.0;  // done
    v.flags[1] = 0.0;  // iterations
    v.flags[2] = 0.0;  // failure
    v.flags[3] = 0.0;  // work-group count of k_cg_reduce_step (as an unsigned int)
  }
}

__global__ void k_cg_setup(const double* __restrict__ B, int nr, int n, const double* small, double* X, double* R,
                           double* P) {
  SmallView v(const_cast<double*>(small), nr);
  const size_t tot = (size_t)nr * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nr);
    const double r = B[q] * (1.0 / v.norms[i]);
    X[q] = 0.0;
    R[q] = r;
    P[q] = r;
  }
}

__global__ void k_cg_rtr(const double* partial, int nblk, int nr, double* small) {
  SmallView v(small, nr);
  reduce_sym(partial, nblk, nr, v.RtR);
}

// LDS of the small steps
struct StepLds {
  double M[kMaxNr * kMaxNr], Q[kMaxNr * kMaxNr], L[kSolveWork], Xs[kMaxNr * kMaxNr];
  int conv;
};

// A = (PtKP)^{-1} RtR with PtKP from the reduced dot partials, on one wave
__device__ __forceinline__ void cg_step_a_wave(const double* partial, int nr, double* small, StepLds& sh) {
  SmallView v(small, nr);
  double rt[kSmallPer];
  gather_nn(v.RtR, rt, nr * nr);  // in flight with the partials' loads
  reduce_sym(partial, 1, nr, sh.M);
  scatter_nn(rt, sh.Q, nr * nr);
  load_nn(sh.M, v.PtKP, nr * nr);
  wave_sync();
  small_solve(sh.M, sh.Q, sh.Xs, nr, sh.L);
  load_nn(sh.Xs, v.A, nr * nr);
}

// convergence test on diag(R R^T) and Psi = (RtR^{-1} RtR2)^T, on one wave
__device__ __forceinline__ void cg_step_b_wave(const double* partial, int nr, double* small, double tolsq,
                                               StepLds& sh) {
  SmallView v(small, nr);
  double rt[kSmallPer];
  gather_nn(v.RtR, rt, nr * nr);  // in flight with the partials' loads
  if (threadIdx.x == 0) sh.conv = 1;
  reduce_sym(partial, 1, nr, sh.Q);  // RtR2
  if ((int)threadIdx.x < nr && !(sh.Q[threadIdx.x * (nr + 1)] < tolsq)) atomicAnd(&sh.conv, 0);
  wave_sync();
  if (sh.conv) {
    if (threadIdx.x == 0) v.flags[0] = 1.0;
    return;
  }
  scatter_nn(rt, sh.M, nr * nr);
  wave_sync();
  small_solve(sh.M, sh.Q, sh.Xs, nr, sh.L);
  load_nn(sh.Xs, v.Psi, nr * nr);
  load_nn(sh.Q, v.RtR, nr * nr);  // R R' = R2 R2' (linop.h:329)
  if (threadIdx.x == 0) v.flags[1] += 1.0;
}

__global__ void __launch_bounds__(64) k_cg_step_a(const double* partial, int nblk, int nr, double* small) {
  (void)nblk;
  SmallView v(small, nr);
  if (v.flags[0] != 0.0) return;
  __shared__ StepLds sh;
  cg_step_a_wave(partial, nr, small, sh);
}

// The dot partials' sum in a fixed order (16 interleaved strands of blocks,
// combined in strand order; a work-group owns 16 of the nr x nr entries),
// then, in the work-group that finishes last, the small step on its first
// wave (STEP 0: A as k_cg_step_a, 1: convergence and Psi as k_cg_step_b):
// one launch instead of two.  The work-groups count themselves on the
// spare flag (reset by k_cg_norms); the last one resets it.  Its release
// fence writes back only the few partial sums this kernel wrote.
template <int STEP>
__global__ void __launch_bounds__(256)
    k_cg_reduce_step(double* partial, int nblk, int nr, double* small, double tolsq) {
  constexpr int kStrands = 16, kEntries = 256 / kStrands;
  SmallView v(small, nr);
  if (v.flags[0] != 0.0) return;  // converged: the same on every work-group
  const int width = nr * nr;
  __shared__ double red[kStrands][kEntries];
  __shared__ StepLds sh;
  __shared__ unsigned last;
  const int e = threadIdx.x % kEntries, g = threadIdx.x / kEntries;
  const int q = blockIdx.x * kEntries + e;
  // a strand's (at most 512 / 16) partials loaded at once, at clamped
  // addresses, and added in block order
  constexpr int kPer = 512 / kStrands;
  double acc = 0.0;
  {
    const int qc = min(q, width - 1);
    double w[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) w[u] = partial[(size_t)min(g + kStrands * u, nblk - 1) * width + qc];
#pragma unroll
    for (int u = 0; u < kPer; u++)
      if (g + kStrands * u < nblk) acc += w[u];
  }
  if (q >= width) acc = 0.0;
  red[g][e] = acc;
  __syncthreads();
  if (g == 0 && q < width) {  // the first 16 lanes of wave 0
    double t = red[0][e];
#pragma unroll
    for (int h = 1; h < kStrands; h++) t += red[h][e];
    partial[q] = t;
  }
  unsigned* ticket = reinterpret_cast<unsigned*>(v.flags + 3);
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // after wave 0's stores of partial[q]
    last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last || threadIdx.x >= 64) return;  // the rest of the work-group ends
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (STEP == 0) cg_step_a_wave(partial, nr, small, sh);
  else cg_step_b_wave(partial, nr, small, tolsq, sh);
}

template <int MODE>  // 0: X += A P, R -= A KP;  1: P = R + Psi P
__global__ void __launch_bounds__(256)
    k_cg_update(int nr, int n, const double* small, double* X, double* R, double* P, const double* KPm) {
  SmallView v(const_cast<double*>(small), nr);
  if (v.flags[0] != 0.0) return;  // converged or failed earlier
  __shared__ double Am[kMaxNr * kMaxNr];
  __shared__ double cp[8][kMaxNr], ck[8][kMaxNr];
  const double* src = MODE == 0 ? v.A : v.Psi;
  for (int q = threadIdx.x; q < nr * nr; q += blockDim.x) Am[q] = src[q];
  __syncthreads();
  const int cl = threadIdx.x / 32, i = threadIdx.x % 32;  // 8 columns x 32 lanes
  for (int base = blockIdx.x * 8; base < n; base += gridDim.x * 8) {
    const int f = base + cl;
    if (f < n) {
      for (int q = i; q < nr; q += 32) {
        cp[cl][q] = P[(size_t)f * nr + q];
        if (MODE == 0) ck[cl][q] = KPm[(size_t)f * nr + q];
      }
    }
    __syncthreads();
    if (f < n) {
      for (int r = i; r < nr; r += 32) {
        double sx = 0.0, sr = 0.0;
        for (int k = 0; k < nr; k++) {
          sx += Am[r + k * nr] * cp[cl][k];
          if (MODE == 0) sr += Am[r + k * nr] * ck[cl][k];
        }
        const size_t q = (size_t)f * nr + r;
        if (MODE == 0) {
          X[q] += sx;
          R[q] -= sr;
        } else {
          P[q] = R[q] + sx;
        }
      }
    }
    __syncthreads();
  }
}

// The same updates at nr = 32 on the FP64 matrix cores: per 16 features a
// wave forms C x^T (32 x 16) with v_mfma_f64_16x16x4f64, C the 32 x 32
// coefficients (A or Psi) in registers as A-operand fragments loaded once.
// Operand layouts (lane l, q = l / 16, j = l % 16): A[i][kk] on lane
// (i = j, kk = q), B[kk][j] on lane (kk = q, j), D[4v + q][j] in slot v.
// The reduction index is mapped k = 8 kk + ks over the 8 MFMA steps ks, so
// a lane's B values are 8 contiguous components of its feature (64 B), and
// the rows of C are permuted (row 4 (i % 4) + i / 4 of a 16-row block on
// fragment row i) so that a lane's 4 output slots are 4 contiguous
// components: every global access is a 16-byte vector.  No LDS.  The sums
// over k run in the matrix core's order, so the result differs from the
// kernel above in the last bits.
template <int MODE>  // 0: X += A P, R -= A KP;  1: P = R + Psi P
__global__ void __launch_bounds__(256)
    k_cg_update32(int n, const double* __restrict__ small, double* __restrict__ X, double* __restrict__ R,
                  double* __restrict__ P, const double* __restrict__ KPm) {
  constexpr int NN = 32 * 32;  // SmallView layout at nr = 32
  if (small[5 * NN + 32] != 0.0) return;  // converged or failed earlier
  const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
  const double* __restrict__ C = small + (MODE == 0 ? 3 : 4) * NN;  // col-major
  double ca[2][8];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int ks = 0; ks < 8; ks++) ca[rb][ks] = C[(16 * rb + 4 * (j & 3) + (j >> 2)) + 32 * (8 * q + ks)];
  const int ntile = (n + 15) / 16;
  const int w = (blockIdx.x * 256 + threadIdx.x) >> 6, nw = (gridDim.x * 256) >> 6;
  for (int t = w; t < ntile; t += nw) {
    const int f = 16 * t + j;
    const bool ok = f < n;
    // every operand of the tile is loaded before the first MFMA, with no
    // branch around the loads (a lane past n reads feature 0: its column of
    // the product is its own and is not stored)
    const size_t o = (size_t)(ok ? f : 0) * 32;
    const dbl4 p0 = ld4(P + o + 8 * q), p1 = ld4(P + o + 8 * q + 4);
    dbl4 k0 = {0.0, 0.0, 0.0, 0.0}, k1 = k0, x0 = k0, x1 = k0;
    if (MODE == 0) {
      k0 = ld4(KPm + o + 8 * q);
      k1 = ld4(KPm + o + 8 * q + 4);
      x0 = ld4(X + o + 4 * q);
      x1 = ld4(X + o + 16 + 4 * q);
    }
    const dbl4 r0 = ld4(R + o + 4 * q), r1 = ld4(R + o + 16 + 4 * q);
    // one wait for all of them: keeps the compiler from sinking the D-layout
    // loads into the store branch (a second round trip per tile)
    asm volatile("" ::"v"(p0), "v"(p1), "v"(r0), "v"(r1));
    if (MODE == 0) asm volatile("" ::"v"(k0), "v"(k1), "v"(x0), "v"(x1));
    auto prod = [&](dbl4 v0, dbl4 v1, dbl4& d0, dbl4& d1) {
      const double b[8] = {v0[0], v0[1], v0[2], v0[3], v1[0], v1[1], v1[2], v1[3]};
      d0 = dbl4{0.0, 0.0, 0.0, 0.0};
      d1 = d0;
#pragma unroll
      for (int ks = 0; ks < 8; ks++) {
        d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[0][ks], b[ks], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[1][ks], b[ks], d1, 0, 0, 0);
      }
    };
    dbl4 d0, d1;
    prod(p0, p1, d0, d1);
    if (MODE == 0) {
      dbl4 e0, e1;
      prod(k0, k1, e0, e1);
      if (ok) {
        st4(X + o + 4 * q, x0 + d0);
        st4(X + o + 16 + 4 * q, x1 + d1);
        st4(R + o + 4 * q, r0 - e0);
        st4(R + o + 16 + 4 * q, r1 - e1);
      }
    } else if (ok) {
      st4(P + o + 4 * q, r0 + d0);
      st4(P + o + 16 + 4 * q, r1 + d1);
    }
  }
}

// X += A P and R -= A KP at nr = 32 as above, fused with the partial sums of
// R^T R over the new R (what the convergence test and Psi need next): a
// block takes `cols` features (a wave a quarter, 16 at a time), so its
// partial has the layout and block count of the dot kernel's.  A wave
// passes its 16 new rows of R through LDS (rows of 34 doubles: 16-byte
// accesses without bank conflicts) from the update's output layout into the
// dot kernel's operand layout: lane (q, j) takes components 2j, 2j + 1 of
// feature 4s + q for the four steps s.  Features past the end enter the
// sums as zero rows.
__global__ void __launch_bounds__(256)
    k_cg_update32_xr(int n, int cols, const double* __restrict__ small, double* __restrict__ X,
                     double* __restrict__ R, const double* __restrict__ P, const double* __restrict__ KPm,
                     double* __restrict__ partial, const int* __restrict__ nxt, const double* __restrict__ reg) {
  constexpr int NN = 32 * 32, LD = 34;
  if (small[5 * NN + 32] != 0.0) return;  // converged or failed earlier
  __shared__ double tile[4][16 * LD];
  __shared__ double red[4][NN];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
  const double* __restrict__ C = small + 3 * NN;  // A, col-major
  double ca[2][8];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int ks = 0; ks < 8; ks++) ca[rb][ks] = C[(16 * rb + 4 * (j & 3) + (j >> 2)) + 32 * (8 * q + ks)];
  const int b0 = blockIdx.x * cols, b1 = min(n, b0 + cols);
  const int per = ((cols + 3) / 4 + 15) / 16 * 16;
  const int c0 = min(b1, b0 + wv * per), c1 = min(b1, c0 + per);
  dbl4 dd[2][2];
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++) dd[u][v] = dbl4{0.0, 0.0, 0.0, 0.0};
  double* T = tile[wv];
  // with nxt, a tile whose features have no nonzero in F^T takes
  // KP = reg * P (not written there); nxt for the next tile loads ahead
  const double lam = nxt ? *reg : 0.0;
  int e = nxt ? nxt[min(c0, n)] : 0;
  for (int f0 = c0; f0 < c1; f0 += 16) {
    const int f = f0 + j;
    const bool ok = f < c1;
    const bool emp = nxt && e >= min(f0 + 16, c1);  // the same on every lane
    if (nxt) e = nxt[min(f0 + 16, n)];
    const size_t o = (size_t)(ok ? f : c0) * 32;
    double bp[8], bk[8];
    const dbl4 p0 = ld4(P + o + 8 * q), p1 = ld4(P + o + 8 * q + 4);
    dbl4 k0, k1;
    if (emp) {
      k0 = lam * p0;
      k1 = lam * p1;
    } else {
      k0 = ld4(KPm + o + 8 * q);
      k1 = ld4(KPm + o + 8 * q + 4);
    }
    const dbl4 x0 = ld4(X + o + 4 * q), x1 = ld4(X + o + 16 + 4 * q);
    const dbl4 r0 = ld4(R + o + 4 * q), r1 = ld4(R + o + 16 + 4 * q);
    // one wait for all eight loads: keeps the compiler from sinking the X and
    // R loads into the store branch (a second round trip per tile)
    asm volatile("" ::"v"(p0), "v"(p1), "v"(k0), "v"(k1));
    asm volatile("" ::"v"(x0), "v"(x1), "v"(r0), "v"(r1));
#pragma unroll
    for (int u = 0; u < 4; u++) {
      bp[u] = p0[u];
      bp[4 + u] = p1[u];
      bk[u] = k0[u];
      bk[4 + u] = k1[u];
    }
    dbl4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0, e0 = d0, e1 = d0;
#pragma unroll
    for (int ks = 0; ks < 8; ks++) {
      d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[0][ks], bp[ks], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[1][ks], bp[ks], d1, 0, 0, 0);
      e0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[0][ks], bk[ks], e0, 0, 0, 0);
      e1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[1][ks], bk[ks], e1, 0, 0, 0);
    }
    dbl4 n0 = r0 - e0, n1 = r1 - e1;
    if (ok) {
      st4(X + o + 4 * q, x0 + d0);
      st4(X + o + 16 + 4 * q, x1 + d1);
      st4(R + o + 4 * q, n0);
      st4(R + o + 16 + 4 * q, n1);
    } else {
      n0 = dbl4{0.0, 0.0, 0.0, 0.0};
      n1 = n0;
    }
    // slot v of lane (q, j): components 4q + v and 16 + 4q + v of feature f0 + j
    st4(T + j * LD + 4 * q, n0);
    st4(T + j * LD + 16 + 4 * q, n1);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // a wave's LDS accesses are performed in order
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const double2 v = *reinterpret_cast<const double2*>(T + (4 * st + q) * LD + 2 * j);
      dd[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, v.x, dd[0][0], 0, 0, 0);
      dd[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, v.y, dd[0][1], 0, 0, 0);
      dd[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, v.x, dd[1][0], 0, 0, 0);
      dd[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, v.y, dd[1][1], 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // the next tile's writes follow these reads
  }
  // slot v of lane (q, j) in tile (ib, jb) is D[2 (4v + q) + ib][2j + jb]
#pragma unroll
  for (int ib = 0; ib < 2; ib++)
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[wv][(2 * (4 * v + q) + ib) + 32 * (2 * j + jb)] = dd[ib][jb][v];
  __syncthreads();
  double* out = partial + (size_t)blockIdx.x * NN;
  for (int e = threadIdx.x; e < NN; e += 256) out[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

__global__ void __launch_bounds__(64) k_cg_step_b(const double* partial, int nblk, int nr, double* small,
                                                   double tolsq) {
  (void)nblk;
  SmallView v(small, nr);
  if (v.flags[0] != 0.0) return;
  __shared__ StepLds sh;
  cg_step_b_wave(partial, nr, small, tolsq, sh);
}

__global__ void k_cg_finish(int nr, int n, const double* small, double* X) {
  SmallView v(const_cast<double*>(small), nr);
  const size_t tot = (size_t)nr * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x)
    X[q] *= v.norms[q % nr];
}

static int grid_for(size_t work, int per) {
  size_t g = (work + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

// the done flag of the CG (SmallView::flags[0]): kernels of iterations after
// convergence return at once
static const double* done_flag(const CgState& c) { return c.small + 5 * c.nr * c.nr + c.nr; }

void launch_cg_init(const CgState& c, const double* B, hipStream_t s) {
  launch_dots(B, B, c.nr, c.n, c.partial, c.nblk, s);
  hipLaunchKernelGGL(k_cg_norms, dim3(1), dim3(64), 0, s, c.partial, 1, c.nr, c.small);
  hipLaunchKernelGGL(k_cg_setup, dim3(grid_for((size_t)c.nr * c.n, 256)), dim3(256), 0, s, B, c.nr, c.n,
                     (const double*)c.small, c.X, c.R, c.P);
  launch_dots(c.R, c.R, c.nr, c.n, c.partial, c.nblk, s);
  hipLaunchKernelGGL(k_cg_rtr, dim3(1), dim3(64), 0, s, c.partial, 1, c.nr, c.small);
}

// more than one partial: the sum and the small step are one launch
static bool fold_steps(const CgState& c) { return c.nblk > 1; }

static void launch_reduce_step(const CgState& c, int step, double tolsq, hipStream_t s) {
  if (c.nblk > 512) throw std::logic_error("k_cg_reduce_step sums at most 512 partials");
  const dim3 grid((c.nr * c.nr + 15) / 16);
  if (step == 0)
    hipLaunchKernelGGL(k_cg_reduce_step<0>, grid, dim3(256), 0, s, c.partial, c.nblk, c.nr, c.small, tolsq);
  else
    hipLaunchKernelGGL(k_cg_reduce_step<1>, grid, dim3(256), 0, s, c.partial, c.nblk, c.nr, c.small, tolsq);
}

void launch_cg_step_a(const CgState& c, hipStream_t s) {
  launch_dots(c.P, c.KPm, c.nr, c.n, c.partial, c.nblk, s, done_flag(c), !fold_steps(c), c.nxt, c.reg);
  if (fold_steps(c)) launch_reduce_step(c, 0, 0.0, s);
  else hipLaunchKernelGGL(k_cg_step_a, dim3(1), dim3(64), 0, s, c.partial, 1, c.nr, c.small);
}

// nr = 32: the update also leaves the reduced partial sums of R^T R, which
// launch_cg_step_b then uses without a dot pass of its own
static bool fused_xr(const CgState& c) { return c.nr == 32; }

void launch_cg_update_xr(const CgState& c, hipStream_t s) {
  if (fused_xr(c)) {
    const int cols = (c.n + c.nblk - 1) / c.nblk;
    hipLaunchKernelGGL(k_cg_update32_xr, dim3(c.nblk), dim3(256), 0, s, c.n, cols, (const double*)c.small, c.X, c.R,
                       (const double*)c.P, (const double*)c.KPm, c.partial, c.nxt, c.reg);
    if (c.nblk > 1 && !fold_steps(c)) launch_partial_reduce(c.partial, c.nblk, 32 * 32, s);
    return;
  }
  hipLaunchKernelGGL((k_cg_update<0>), dim3(grid_for(c.n, 8)), dim3(256), 0, s, c.nr, c.n, (const double*)c.small,
                     c.X, c.R, c.P, (const double*)c.KPm);
}

void launch_cg_step_b(const CgState& c, double tolsq, hipStream_t s) {
  if (!fused_xr(c)) launch_dots(c.R, c.R, c.nr, c.n, c.partial, c.nblk, s, done_flag(c), !fold_steps(c));
  if (fold_steps(c)) launch_reduce_step(c, 1, tolsq, s);
  else hipLaunchKernelGGL(k_cg_step_b, dim3(1), dim3(64), 0, s, c.partial, 1, c.nr, c.small, tolsq);
}

void launch_cg_update_p(const CgState& c, hipStream_t s) {
  if (c.nr == 32) {
    hipLaunchKernelGGL((k_cg_update32<1>), dim3((c.n + 63) / 64), dim3(256), 0, s, c.n, (const double*)c.small, c.X,
                       c.R, c.P, (const double*)c.KPm);
    return;
  }
  hipLaunchKernelGGL((k_cg_update<1>), dim3(grid_for(c.n, 8)), dim3(256), 0, s, c.nr, c.n, (const double*)c.small,
                     c.X, c.R, c.P, (const double*)c.KPm);
}

void launch_cg_finish(const CgState& c, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_finish, dim3(grid_for((size_t)c.nr * c.n, 256)), dim3(256), 0, s, c.nr, c.n,
                     (const double*)c.small, c.X);
}

// ------------------------------------------------------------------
// sample_beta (latentprior.cpp:272-293) draws, wave per column:
//   tmp[:,n]  = U[:,n] + chol(Lambda)^{-T} z_n - mu
//   Fty[:,f] += sqrt(lambda_beta) chol(Lambda)^{-T} z'_f
// MvNormal_prec_omp (mvnormal.cpp:186-197).
// ------------------------------------------------------------------
__device__ void wave_back_solve(const double* __restrict__ L, int K, double x[2], int lane) {
  // x holds components lane and lane+64; solves L^T x = z in place
  for (int j = K - 1; j >= 0; j--) {
    const int owner = j & 63, slot = j >> 6;
    double xj = __shfl(slot ? x[1] : x[0], owner, 64);
    xj = xj / L[j + j * K];
    if (lane == owner) x[slot] = xj;
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      if (i < j) x[s] -= L[j + i * K] * xj;
    }
  }
}

template <int WHICH>  // 0: U columns, 1: feature columns
__global__ void __launch_bounds__(256) k_beta_draw(BetaArgs a, double* out) {
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int nw = (gridDim.x * 256) >> 6;
  const int it = *a.iter;
  const int ncol = WHICH == 0 ? a.N : a.nfeat;
  const uint32_t tg = make_tag(WHICH == 0 ? TAG_BETA_U : TAG_BETA_F, a.mode);
  const double sl = WHICH == 1 ? sqrt(*a.lambda_beta) : 0.0;
  for (int c = w; c < ncol; c += nw) {
    double x[2];
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      x[s] = i < a.K ? normal_f64(a.seed, (uint32_t)c, (uint32_t)i, (uint32_t)it, tg) : 0.0;
    }
    wave_back_solve(a.Lchol, a.K, x, lane);
    for (int s = 0; s < 2; s++) {
      const int i = lane + 64 * s;
      if (i < a.K) {
        const size_t q = (size_t)c * a.K + i;
        if (WHICH == 0) out[q] = (double)a.U[(size_t)c * a.KP + i] + x[s] - a.mu[i];
        else out[q] += sl * x[s];
      }
    }
  }
}

// The same draws for K <= 32, one column per thread: the K normals of a
// column come four per Philox block, and L^T x = z is solved in registers
// with chol(Lambda) and its reciprocal diagonal read as LDS broadcasts.
template <int WHICH, int KP>
__global__ void __launch_bounds__(256) k_beta_draw_t(BetaArgs a, double* out) {
  __shared__ double Ls[KP * KP], inv[KP];
  const int K = a.K;
  for (int q = threadIdx.x; q < KP * KP; q += 256) {
    const int i = q % KP, j = q / KP;
    Ls[q] = (i < K && j < K) ? a.Lchol[i + j * K] : 0.0;
  }
  for (int q = threadIdx.x; q < KP; q += 256) inv[q] = q < K ? 1.0 / a.Lchol[q + q * K] : 0.0;
  __syncthreads();
  const int it = *a.iter;
  const int ncol = WHICH == 0 ? a.N : a.nfeat;
  const uint32_t tg = make_tag(WHICH == 0 ? TAG_BETA_U : TAG_BETA_F, a.mode);
  const double sl = WHICH == 1 ? sqrt(*a.lambda_beta) : 0.0;
  for (int c = blockIdx.x * 256 + threadIdx.x; c < ncol; c += gridDim.x * 256) {
    double x[KP];
#pragma unroll
    for (int b = 0; b < KP / 4; b++) {
      // one Philox block at a time (the column index is passed through an
      // opaque dependency on the previous block, which bounds live registers)
      int cb = c;
      opaque_after(cb, b > 0 ? x[4 * b - 1] : 0.0);
      double z[4];
      normal_f64x4(a.seed, (uint32_t)cb, (uint32_t)b, (uint32_t)it, tg, z);
#pragma unroll
      for (int u = 0; u < 4; u++) x[4 * b + u] = 4 * b + u < K ? z[u] : 0.0;
    }
    // L^T x = z from the last row up: x_j = (z_j - sum_{i>j} L[i][j] x_i) / L[j][j]
#pragma unroll
    for (int j = KP - 1; j >= 0; j--) {
      int o = 0;
      opaque_after(o, j < KP - 1 ? x[j + 1] : 0.0);
      double acc = x[j];
#pragma unroll
      for (int i = j + 1; i < KP; i++) acc -= Ls[i + j * KP + o] * x[i];
      x[j] = acc * inv[j + o];
    }
    if (WHICH == 0) {
      const float* u = a.U + (size_t)c * a.KP;
#pragma unroll
      for (int i = 0; i < KP; i++)
        if (i < K) out[(size_t)c * K + i] = (double)u[i] + x[i] - a.mu[i];
    } else {
#pragma unroll
      for (int i = 0; i < KP; i++)
        if (i < K) out[(size_t)c * K + i] += sl * x[i];
    }
  }
}

template <int WHICH>
static void launch_beta_draw(const BetaArgs& a, double* out, int ncol, hipStream_t s) {
  if (ncol <= 0) return;
  const int KP = pad_latent(a.K);
  const int g = grid_for(ncol, 256);
  switch (KP) {
    case 8: hipLaunchKernelGGL((k_beta_draw_t<WHICH, 8>), dim3(g), dim3(256), 0, s, a, out); return;
    case 16: hipLaunchKernelGGL((k_beta_draw_t<WHICH, 16>), dim3(g), dim3(256), 0, s, a, out); return;
    case 32: hipLaunchKernelGGL((k_beta_draw_t<WHICH, 32>), dim3(g), dim3(256), 0, s, a, out); return;
    default: hipLaunchKernelGGL((k_beta_draw<WHICH>), dim3(grid_for(ncol, 4)), dim3(256), 0, s, a, out); return;
  }
}

void launch_beta_tmp(const BetaArgs& a, double* tmp, hipStream_t s) { launch_beta_draw<0>(a, tmp, a.N, s); }
void launch_beta_noise_add(const BetaArgs& a, double* Fty, hipStream_t s) {
  launch_beta_draw<1>(a, Fty, a.nfeat, s);
}

// Uhat[:,n] = sum_f F[n,f] beta[:,f] (compute_uhat, linop.h:196-198) as an
// SpMM into FP64 (tmp, K x N), then the per-row prior term
// Lambda (mu + Uhat_n) used by MacauPrior::sample_latents
// (latentprior.cpp:218-227), one thread per output entry.
__global__ void __launch_bounds__(256)
    k_uhat_post(int N, int K, int KP, const double* __restrict__ uh, const double* __restrict__ mu,
                const double* __restrict__ Lambda, float* __restrict__ Uhat, float* __restrict__ Lmu_rows) {
  const size_t tot = (size_t)N * KP;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < tot; q += (size_t)gridDim.x * 256) {
    const int r = (int)(q / KP), i = (int)(q % KP);
    float u = 0.f, lm = 0.f;
    if (i < K) {
      const double* col = uh + (size_t)r * K;
      double t = 0.0;
      for (int j = 0; j < K; j++) t += Lambda[i + j * K] * (mu[j] + col[j]);
      u = (float)col[i];
      lm = (float)t;
    }
    Uhat[q] = u;
    Lmu_rows[q] = lm;
  }
}

void launch_uhat_finish(const CsrDev& F, const double* beta, int K, int KP, const double* mu, const double* Lambda,
                        float* Uhat, float* Lmu_rows, double* tmp, hipStream_t s) {
  if (F.nrow == 0) return;
  launch_spmm(F, beta, K, tmp, nullptr, nullptr, s);
  hipLaunchKernelGGL(k_uhat_post, dim3(grid_for((size_t)F.nrow * KP, 256)), dim3(256), 0, s, F.nrow, K, KP,
                     (const double*)tmp, mu, Lambda, Uhat, Lmu_rows);
}

// lambda_beta ~ Gamma(b, c) with posterior_lambda_beta (latentprior.cpp:317-331),
// nu0 = 1e-3, mu0 = 1 (latentprior.cpp:212-215)
__global__ void k_lambda_beta(int K, double nfeat, const double* bbt_partial, int nblk, const double* Lambda,
                              double* lambda_beta, uint64_t seed, const int* iter, int mode) {
  __shared__ double red[256];
  const int KK = K * K;
  double tr = 0.0;
  for (int q = threadIdx.x; q < KK; q += 256) {
    const int i = q % K, j = q / K;
    const int src = (i >= j) ? q : (j + i * K);
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += bbt_partial[(size_t)b * (K + KK) + K + src];
    tr += s * Lambda[j + i * K];
  }
  red[threadIdx.x] = tr;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nu = 1e-3, mu = 1.0;
    const double nux = nu + (double)K * nfeat;
    const double mux = mu * nux / (nu + mu * red[0]);
    const double b = nux / 2.0, c = 2.0 * mux / nux;
    *lambda_beta = gamma_f64(seed, b, 0u, (uint32_t)*iter, make_tag(TAG_LAMBDA_BETA, mode)) * c;
  }
}

void launch_lambda_beta(int K, double nfeat, const double* bbt_partial, int nblk, const double* Lambda,
                        double* lambda_beta, uint64_t seed, const int* iter, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_lambda_beta, dim3(1), dim3(256), 0, s, K, nfeat, bbt_partial, nblk, Lambda, lambda_beta,
                     seed, iter, mode);
}

}  // namespace mgpu

namespace mgpu {

__global__ void k_rows_copy(const double* __restrict__ src, int lds, double* __restrict__ dst, int ldd, int r0s,
                            int r0d, int nr, int n) {
  const size_t tot = (size_t)nr * n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nr);
    const size_t c = q / nr;
    dst[c * ldd + r0d + i] = src[c * lds + r0s + i];
  }
}

void launch_rows_copy(const double* src, int lds, double* dst, int ldd, int r0s, int r0d, int nr, int n,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_rows_copy, dim3(grid_for((size_t)nr * n, 256)), dim3(256), 0, s, src, lds, dst, ldd, r0s, r0d,
                     nr, n);
}

template <typename T>
__global__ void k_sumsq(const T* __restrict__ x, size_t n, double* partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
    const double v = (double)x[q];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

void launch_sumsq_f32(const float* x, size_t n, double* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_sumsq<float>), dim3(nblk), dim3(256), 0, s, x, n, partial);
}
void launch_sumsq_f64(const double* x, size_t n, double* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL((k_sumsq<double>), dim3(nblk), dim3(256), 0, s, x, n, partial);
}

}  // namespace mgpu
