// model.cpp -- Gibbs driver, data layer, priors and noise models.
//
// Control flow follows MacauX::run (macau.cpp:94-148): per iteration,
// sample every mode's latents, then update every prior, then the noise
// precision, then evaluate.  All state stays on the GPU; one iteration is a
// fixed sequence of kernel launches on one stream, captured once into a HIP
// graph and replayed when no host round trip is needed (no link-matrix CG).
#include "model.h"

#include <signal.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <numeric>
#include <sstream>

namespace mgpu {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " in " + what);
  }
}

void DevBuf::alloc(size_t bytes) {
  release();
  if (bytes == 0) bytes = 16;
  HIPCHK(hipMalloc(&p_, bytes));
  n_ = bytes;
}
void DevBuf::release() {
  if (p_) (void)hipFree(p_);
  p_ = nullptr;
  n_ = 0;
}

template <typename T>
static void upload_scalar(DevBuf& b, T v, hipStream_t s) {
  if (!b.as<void>()) b.alloc(sizeof(T));
  HIPCHK(hipMemcpyAsync(b.as<void>(), &v, sizeof(T), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
}
template <typename T>
static T download_scalar(const DevBuf& b, hipStream_t s) {
  T v;
  HIPCHK(hipMemcpyAsync(&v, b.as<void>(), sizeof(T), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return v;
}

static double tick() {
  return std::chrono::duration<double>(std::chrono::high_resolution_clock::now().time_since_epoch()).count();
}

static std::string to_string_with_precision(double v, int n) {  // bpmfutils.h:175-180
  std::ostringstream out;
  out << std::setprecision(n) << v;
  return out.str();
}

// writeToCSVfile (bpmfutils.h:169-173): 6 significant digits, rows by "\n"
static void write_csv(const std::string& fn, const std::vector<double>& colmajor, int rows, long long cols) {
  std::ofstream f(fn.c_str());
  f << std::setprecision(6);
  for (int i = 0; i < rows; i++) {
    for (long long j = 0; j < cols; j++) {
      f << colmajor[(size_t)j * rows + i];
      if (j + 1 < cols) f << ",";
    }
    if (i + 1 < rows) f << "\n";
  }
}

// ------------------------------------------------------------------
// data layer (sparsetensor.cpp:11-60, 78-117; bpmfutils.h:57-82)
// ------------------------------------------------------------------
struct Cell {
  long long key;
  int pos;
};

// Eigen setFromTriplets: duplicates summed; returns cells sorted by
// (outer, inner) with outer = column (CSC).
static void sum_duplicates(const int* idx, int nnz, const double* vals, const int* dims, std::vector<int>& r,
                           std::vector<int>& c, std::vector<double>& v) {
  std::vector<Cell> cells(nnz);
  for (int e = 0; e < nnz; e++) {
    const int row = idx[e], col = idx[nnz + e];
    if (row < 0 || row >= dims[0] || col < 0 || col >= dims[1])
      throw std::runtime_error("MatrixData: index out of range");
    cells[e] = {(long long)col * dims[0] + row, e};
  }
  std::stable_sort(cells.begin(), cells.end(), [](const Cell& a, const Cell& b) { return a.key < b.key; });
  r.clear();
  c.clear();
  v.clear();
  for (int e = 0; e < nnz; e++) {
    if (e > 0 && cells[e].key == cells[e - 1].key) {
      v.back() += vals[cells[e].pos];
      continue;
    }
    r.push_back((int)(cells[e].key % dims[0]));
    c.push_back((int)(cells[e].key / dims[0]));
    v.push_back(vals[cells[e].pos]);
  }
}

// group cells by key (stable), keeping the other indices in view order
static HostRows group_rows(const std::vector<int>& key, int nkeys, const std::vector<int>& other, int nother,
                           const std::vector<double>& vals) {
  HostRows h;
  const size_t n = key.size();
  h.row_ptr.assign(nkeys + 1, 0);
  for (size_t e = 0; e < n; e++) h.row_ptr[key[e] + 1]++;
  for (int i = 0; i < nkeys; i++) h.row_ptr[i + 1] += h.row_ptr[i];
  std::vector<int> pos(h.row_ptr.begin(), h.row_ptr.end() - 1);
  h.idx.resize(n * nother);
  h.vals.resize(n);
  for (size_t e = 0; e < n; e++) {
    const int d = pos[key[e]]++;
    for (int t = 0; t < nother; t++) h.idx[(size_t)d * nother + t] = other[e * nother + t];
    h.vals[d] = vals[e];
  }
  return h;
}

void Model::setRelationData(const int* idx, int nm, const double* vals, int nnz_in, const int* d) {
  if (nm != nmodes) throw std::runtime_error("setRelationData: number of modes differs from the model");
  dims.assign(d, d + nm);
  rows.clear();
  if (!is_tensor) {
    if (nm != 2) throw std::runtime_error("MatrixData: tensor training input not supported.");
    std::vector<int> r, c;
    std::vector<double> v;
    sum_duplicates(idx, nnz_in, vals, d, r, c, v);
    nnz = (long long)v.size();
    double s = 0.0;
    for (double x : v) s += x;
    mean_value = s / (double)nnz;  // Y.sum() / Y.nonZeros() (sparsetensor.cpp:103)
    // mode 0: rows of Y, column indices ascending (Yt columns)
    std::vector<size_t> order(r.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return r[a] < r[b]; });
    std::vector<int> k0(r.size()), o0(r.size());
    std::vector<double> v0(r.size());
    for (size_t q = 0; q < order.size(); q++) {
      k0[q] = r[order[q]];
      o0[q] = c[order[q]];
      v0[q] = v[order[q]];
    }
    rows.push_back(group_rows(k0, d[0], o0, 1, v0));
    // mode 1: columns of Y, row indices ascending (Y columns)
    rows.push_back(group_rows(c, d[1], r, 1, v));
  } else {
    for (int e = 0; e < nnz_in; e++)
      for (int m = 0; m < nm; m++)
        if (idx[(size_t)m * nnz_in + e] < 0 || idx[(size_t)m * nnz_in + e] >= d[m])
          throw std::runtime_error("SparseMode: mode value larger than mode_size");
    nnz = nnz_in;
    double s = 0.0;
    for (int e = 0; e < nnz_in; e++) s += vals[e];
    mean_value = s / (double)nnz_in;  // vals.mean() (sparsetensor.cpp:127)
    std::vector<double> v(vals, vals + nnz_in);
    for (int m = 0; m < nm; m++) {
      std::vector<int> key(nnz_in), other((size_t)nnz_in * (nm - 1));
      for (int e = 0; e < nnz_in; e++) {
        key[e] = idx[(size_t)m * nnz_in + e];
        int t = 0;
        for (int q = 0; q < nm; q++)
          if (q != m) other[(size_t)e * (nm - 1) + t++] = idx[(size_t)q * nnz_in + e];
      }
      rows.push_back(group_rows(key, d[m], other, nm - 1, v));
    }
  }
}

void Model::setRelationDataTest(const int* idx, int nm, const double* vals, int nnz_in, const int* d) {
  if (nm != nmodes) throw std::runtime_error("setRelationDataTest: number of modes differs from the model");
  has_test = true;
  test_idx.clear();
  test_vals.clear();
  if (!is_tensor) {
    std::vector<int> r, c;
    std::vector<double> v;
    sum_duplicates(idx, nnz_in, vals, d, r, c, v);  // Ytest in CSC order
    for (size_t q = 0; q < v.size(); q++) {
      test_idx.push_back(r[q]);
      test_idx.push_back(c[q]);
    }
    test_vals = v;
  } else {
    for (int m = 0; m < nm; m++)
      if (!dims.empty() && d[m] != dims[m])
        throw std::runtime_error("setTest(): train and test Tensor sizes are not equal.");
    for (int e = 0; e < nnz_in; e++)
      for (int m = 0; m < nm; m++)
        if (idx[(size_t)m * nnz_in + e] < 0 || idx[(size_t)m * nnz_in + e] >= d[m])
          throw std::runtime_error("SparseMode: mode value larger than mode_size");
    std::vector<int> order(nnz_in);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return idx[a] < idx[b]; });
    for (int q : order) {
      for (int m = 0; m < nm; m++) test_idx.push_back(idx[(size_t)m * nnz_in + q]);
      test_vals.push_back(vals[q]);
    }
  }
  ntest = (long long)test_vals.size();
}

// ------------------------------------------------------------------
// Model
// ------------------------------------------------------------------
Model::Model(int nm, int K, std::unique_ptr<INoiseModel> nz)
    : nmodes(nm), num_latent(K), KP(pad_latent(K)), noise(std::move(nz)) {
  is_tensor = nm != 2;
  if (K < 1 || K > 128) throw std::runtime_error("num_latent must be in [1, 128]");
  if (nm < 2 || nm > kMaxOther + 1) throw std::runtime_error("number of modes must be in [2, 10]");
}

void Model::drop_graphs() {
  for (int v = 0; v < 2; v++) {
    if (graph_exec[v]) (void)hipGraphExecDestroy(graph_exec[v]);
    if (graph[v]) (void)hipGraphDestroy(graph[v]);
    graph_exec[v] = nullptr;
    graph[v] = nullptr;
  }
}

Model::~Model() {
  drop_graphs();
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (side) {
    (void)hipStreamSynchronize(side);
    (void)hipStreamDestroy(side);
  }
  if (seg_stream) {
    (void)hipStreamSynchronize(seg_stream);
    (void)hipStreamDestroy(seg_stream);
  }
  for (auto& e : ev_fork) (void)hipEventDestroy(e);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (ev_start) (void)hipEventDestroy(ev_start);
  if (ev_plast) (void)hipEventDestroy(ev_plast);
  if (ev_noise) (void)hipEventDestroy(ev_noise);
  for (auto& e : ev_seg_fork) (void)hipEventDestroy(e);
  for (auto& e : ev_seg_join) (void)hipEventDestroy(e);
}

FactorSet Model::others_of(int mode) const {
  FactorSet f;
  int t = 0;
  for (int q = 0; q < nmodes; q++)
    if (q != mode) {
      f.n[t] = dims[q];
      f.f[t++] = d_U[q].as<float>();
    }
  return f;
}

void Model::init() {
  if ((int)priors.size() < 2) throw std::runtime_error("Less 2 priors are not supported.");
  if ((int)priors.size() != nmodes)
    throw std::runtime_error("number of priors must equal the number of data modes");
  if (rows.empty()) throw std::runtime_error("training data not set");
  KP = sampler == SAMPLER_BATCHED ? pad_latent_batched(num_latent) : pad_latent(num_latent);
  if (!stream) HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (!side) HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  if (!seg_stream) HIPCHK(hipStreamCreateWithFlags(&seg_stream, hipStreamNonBlocking));
  if (ev_fork.empty()) {
    ev_fork.resize(nmodes);
    for (auto& e : ev_fork) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_start, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_plast, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_noise, hipEventDisableTiming));
    ev_seg_fork.resize(nmodes);
    ev_seg_join.resize(nmodes);
    for (auto& e : ev_seg_fork) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : ev_seg_join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  drop_graphs();
  prior_pending = false;
  d_rowptr.clear(); d_idx.clear(); d_r.clear(); d_U.clear(); rdev.clear();
  for (int m = 0; m < nmodes; m++) {
    const HostRows& h = rows[m];
    std::vector<float> r(h.vals.size());
    for (size_t e = 0; e < r.size(); e++) r[e] = (float)(h.vals[e] - mean_value);
    d_rowptr.emplace_back();
    d_idx.emplace_back();
    d_r.emplace_back();
    upload(d_rowptr.back(), h.row_ptr, stream);
    upload(d_idx.back(), h.idx, stream);
    upload(d_r.back(), r, stream);
    d_U.emplace_back((size_t)KP * dims[m] * sizeof(float));
    HIPCHK(hipMemsetAsync(d_U.back().as<void>(), 0, (size_t)KP * dims[m] * sizeof(float), stream));
    RowsDev rd;
    rd.N = dims[m];
    rd.nother = nmodes - 1;
    rd.row_ptr = d_rowptr.back().as<int>();
    rd.idx = d_idx.back().as<int>();
    rd.r = d_r.back().as<float>();
    rdev.push_back(rd);
    HIPCHK(hipStreamSynchronize(stream));  // host vectors r go out of scope
  }
  upload_scalar<int>(d_iter, 0, stream);
  upload_scalar<int>(d_iter_prior, 0, stream);
  upload_scalar<int>(d_fail, 0, stream);
  d_eval_out.alloc(2 * sizeof(double));
  d_sse.alloc(sizeof(double));
  d_alpha.alloc(sizeof(double));
  // test cells, on the device in the order of the largest mode's index so
  // that consecutive cells share that mode's factor row (L2 hits in the
  // evaluation's gathers); d_pred and d_pred_var follow this order and
  // test_to_host restores the caller's
  {
    int km = 0;
    for (int m = 1; m < nmodes; m++)
      if (dims[m] > dims[km]) km = m;
    test_perm.resize(ntest);
    std::iota(test_perm.begin(), test_perm.end(), 0LL);
    std::stable_sort(test_perm.begin(), test_perm.end(), [&](long long a, long long b) {
      return test_idx[(size_t)a * nmodes + km] < test_idx[(size_t)b * nmodes + km];
    });
    std::vector<int> didx((size_t)ntest * nmodes);
    std::vector<double> dvals(ntest);
    for (long long t = 0; t < ntest; t++) {
      const long long q = test_perm[t];
      for (int m = 0; m < nmodes; m++) didx[(size_t)t * nmodes + m] = test_idx[(size_t)q * nmodes + m];
      dvals[t] = test_vals[q];
    }
    upload(d_test_idx, didx, stream);
    upload(d_test_vals, dvals, stream);
    HIPCHK(hipStreamSynchronize(stream));  // didx and dvals go out of scope
  }
  d_pred.alloc(std::max<long long>(ntest, 1) * sizeof(double));
  d_pred_var.alloc(std::max<long long>(ntest, 1) * sizeof(double));
  HIPCHK(hipMemsetAsync(d_pred.as<void>(), 0, d_pred.bytes(), stream));
  HIPCHK(hipMemsetAsync(d_pred_var.as<void>(), 0, d_pred_var.bytes(), stream));
  eval_nblk = eval_blocks((int)ntest);
  d_eval_partial.alloc((size_t)2 * eval_nblk * sizeof(double));
  noise_nblk = noise_sse_blocks(dims[0]);
  d_noise_partial.alloc((size_t)noise_nblk * sizeof(double));
  d_norm_partial.alloc(256 * sizeof(double));
  for (int m = 0; m < nmodes; m++) priors[m]->init(*this, m);
  noise->init(*this);
  apply_cached_tuning();
  build_plans();
  rq = RefineQueue();
  refine_bufs.clear();
  if (sampler == SAMPLER_BATCHED && refine && KP <= 64) {  // no refinement step for K > 64
    const int maxrows = *std::max_element(dims.begin(), dims.end());
    rq = make_refine_queue(refine_bufs, KP, std::min(maxrows, kRefineCap), refine_thresh,
                           refine_min_nnz < 0 ? std::max(num_latent, kRefineMinNnz) : refine_min_nnz, stream);
  }
  // the batched sampler computes the train SSE of the adaptive noise inside
  // the last half-sweep (u^T G u - 2 u^T c + sum r^2 per row)
  // (not when the last mode is univariate: that mode runs the plain sampler)
  fused_sse = sampler == SAMPLER_BATCHED && dynamic_cast<AdaptiveGaussianNoise*>(noise.get()) != nullptr &&
              dynamic_cast<MacauOnePrior*>(priors[nmodes - 1].get()) == nullptr;
  if (fused_sse) {
    const int last = nmodes - 1;
    const HostRows& h = rows[last];
    std::vector<float> r2(dims[last]);
    for (int n = 0; n < dims[last]; n++) {
      double s = 0.0;
      for (int e = h.row_ptr[n]; e < h.row_ptr[n + 1]; e++) {
        const double r = (double)(float)(h.vals[e] - mean_value);
        s += r * r;
      }
      r2[n] = (float)s;
    }
    upload(d_r2_last, r2, stream);
    // slots for the largest launch shape the autotuner may choose
    d_sse_partial.alloc((size_t)(kMaxLightCap + kMaxHeavyCap) * sizeof(double));
    update_sse_slots();
  }
  HIPCHK(hipStreamSynchronize(stream));
  host_iter = 0;
  initialized = true;
}

RefineQueue make_refine_queue(std::vector<DevBuf>& B, int KP, int cap, float thresh, int min_nnz, hipStream_t s) {
  B.clear();
  B.resize(6);
  upload(B[0], std::vector<int>{0, 0}, s);
  B[1].alloc((size_t)std::max(cap, 1) * sizeof(int));
  B[2].alloc((size_t)std::max(cap, 1) * KP * KP * sizeof(float));
  upload(B[3], std::vector<long long>{0, 0}, s);
  B[4].alloc((size_t)kRefineGrid * KP * sizeof(double));
  upload(B[5], std::vector<int>(kRefineGrid, 0), s);
  RefineQueue q;
  q.count = B[0].as<int>();
  q.cap = cap;
  q.row = B[1].as<int>();
  q.L = B[2].as<float>();
  q.thresh = thresh;
  q.min_nnz = min_nnz;
  q.totals = B[3].as<long long>();
  q.part = B[4].as<double>();
  q.row_ticket = B[5].as<int>();
  return q;
}

void Model::set_iteration(int it) {
  launch_pending_prior();
  upload_scalar<int>(d_iter, it, stream);
  upload_scalar<int>(d_iter_prior, it, stream);
  host_iter = it;
}

BatchPlan build_batch_plan(const std::vector<int>& row_ptr, int KP, std::vector<DevBuf>& B, hipStream_t s,
                           int seg_len) {
  const int N = (int)row_ptr.size() - 1;
  // rows sorted by nnz (descending); rows above the segment length are split
  // (default 2048: fewer, longer segments measured faster than 1024 and 512)
  if (seg_len < 64 || seg_len % 64) throw std::runtime_error("segment length must be a positive multiple of 64");
  const int seg = seg_len;
  std::vector<int> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (row_ptr[a + 1] - row_ptr[a]) > (row_ptr[b + 1] - row_ptr[b]);
  });
  std::vector<int> light, heavy, hseg0, sr, sb, se;
  for (int q = 0; q < N; q++) {
    const int r = order[q];
    const int nz = row_ptr[r + 1] - row_ptr[r];
    if (nz > seg) {
      hseg0.push_back((int)sb.size());
      for (int b = row_ptr[r]; b < row_ptr[r + 1]; b += seg) {
        sr.push_back((int)heavy.size());
        sb.push_back(b);
        se.push_back(std::min(b + seg, row_ptr[r + 1]));
      }
      heavy.push_back(r);
    } else {
      light.push_back(r);
    }
  }
  hseg0.push_back((int)sb.size());
  // the light rows as the finishing kernel takes them (pairs for KP = 32)
  const int per = KP == 32 ? 2 : 1;
  const int npair = ((int)light.size() + per - 1) / per;
  std::vector<int> prng(4 * (size_t)npair, 0), prow(2 * (size_t)npair, -1);
  for (int q = 0; q < npair; q++) {
    for (int u = 0; u < per && q * per + u < (int)light.size(); u++) {
      const int r = light[q * per + u];
      prow[2 * q + u] = r;
      prng[4 * q + 2 * u] = row_ptr[r];
      prng[4 * q + 2 * u + 1] = row_ptr[r + 1];
    }
  }
  B.clear();
  B.resize(9);
  upload(B[7], prng, s);
  upload(B[8], prow, s);
  upload(B[0], light, s);
  upload(B[1], sr, s);
  upload(B[2], sb, s);
  upload(B[3], se, s);
  upload(B[4], heavy, s);
  upload(B[5], hseg0, s);
  // per segment: the KP x KP Gram fragment and one RHS partial per lane (64)
  B[6].alloc(std::max<size_t>(sb.size(), 1) * (size_t)(KP * KP + 64) * sizeof(float));
  BatchPlan p;
  p.KP = KP;
  p.n_light = (int)light.size();
  p.light = B[0].as<int>();
  p.n_seg = (int)sb.size();
  p.seg_row = B[1].as<int>();
  p.seg_beg = B[2].as<int>();
  p.seg_end = B[3].as<int>();
  p.n_heavy = (int)heavy.size();
  p.heavy = B[4].as<int>();
  p.heavy_seg0 = B[5].as<int>();
  p.partial = B[6].as<float>();
  p.n_pair = npair;
  p.pair_rng = reinterpret_cast<const int4*>(B[7].as<int>());
  p.pair_row = reinterpret_cast<const int2*>(B[8].as<int>());
  p.seg_len = seg;
  HIPCHK(hipStreamSynchronize(s));
  return p;
}

void Model::build_plans() {
  plans.assign(nmodes, BatchPlan());
  plan_bufs.clear();
  plan_bufs.resize(nmodes);
  if (sampler != SAMPLER_BATCHED) return;
  for (int m = 0; m < nmodes; m++) {
    if (!batched_supported(KP, nmodes - 1))
      throw std::runtime_error("batched sampler: unsupported latent size / mode count");
    if (!batched_fits(others_of(m), nmodes - 1, KP))
      throw std::runtime_error("batched sampler: a factor matrix exceeds 2^31 bytes");
    plans[m] = build_batch_plan(rows[m].row_ptr, KP, plan_bufs[m], stream, tune[m].seg_len);
    apply_plan_shape(m);
  }
}

SampleArgs Model::sample_args(int mode, const ILatentPrior& p) {
  SampleArgs a;
  a.rows = rdev[mode];
  a.others = others_of(mode);
  a.alpha = d_alpha.as<double>();
  a.Lambda_f = p.Lambda_f.as<float>();
  a.Lmu_f = p.Lmu_f.as<float>();
  a.Lmu_rows = p.prior_mean_rows();
  a.U = d_U[mode].as<float>();
  a.K = num_latent;
  a.KP = KP;
  a.seed = seed;
  a.iter = d_iter.as<int>();
  a.mode = mode;
  a.fail = d_fail.as<int>();
  if (fused_sse && mode == nmodes - 1) {
    a.sse_partial = d_sse_partial.as<double>();
    a.r2 = d_r2_last.as<float>();
  }
  if (sampler == SAMPLER_BATCHED && refine && KP <= 64) a.refine = rq;
  return a;
}

namespace {
// makes Model::stream the side stream for the lifetime of the guard, so the
// prior code (which launches on m.stream) runs there
struct OnSideStream {
  Model& m;
  explicit OnSideStream(Model& mm) : m(mm) { std::swap(m.stream, m.side); }
  ~OnSideStream() { std::swap(m.stream, m.side); }
};
}  // namespace

void Model::step_launch(bool defer_last, bool run_pending) {
  // The reference samples every mode, then updates every prior
  // (macau.cpp:120-128).  The update of mode m reads only mode m's new
  // samples and is read only by mode m's next sampling; its draws are keyed
  // by (iteration, mode) through d_iter_prior.  So it is launched on the side
  // stream as soon as mode m is sampled and runs while the next modes are
  // sampled.  The last mode's update either runs beside the noise and the
  // test predictions, or, with defer_last, is left for the next launch,
  // where it runs beside the first modes' sampling and the last mode waits
  // for it (run_pending): an iteration then ends with its evaluation instead
  // of the last Normal-Wishart draw.  Same data dependencies, same draws,
  // same results.
  const int last = nmodes - 1;
  if (run_pending) {
    HIPCHK(hipEventRecord(ev_start, stream));
    HIPCHK(hipStreamWaitEvent(side, ev_start, 0));
    {
      OnSideStream guard(*this);
      priors[last]->update_prior(*this, last);
    }
    launch_advance_iter(d_iter_prior.as<int>(), side);
    HIPCHK(hipEventRecord(ev_plast, side));
  }
  for (int m = 0; m < nmodes; m++) {
    if (m == last && run_pending) HIPCHK(hipStreamWaitEvent(stream, ev_plast, 0));
    noise->augment(*this, m);
    priors[m]->sample_latents(*this, m);
    if (m == last && defer_last) break;
    HIPCHK(hipEventRecord(ev_fork[m], stream));
    HIPCHK(hipStreamWaitEvent(side, ev_fork[m], 0));
    OnSideStream guard(*this);
    priors[m]->update_prior(*this, m);
    if (m == last) launch_advance_iter(d_iter_prior.as<int>(), stream);  // (the side stream, swapped in)
  }
  // the noise precision reads only the sampled factors (or the SSE the last
  // half-sweep fused): drawn on the side stream beside the evaluation
  HIPCHK(hipEventRecord(ev_noise, stream));
  HIPCHK(hipStreamWaitEvent(side, ev_noise, 0));
  {
    OnSideStream guard(*this);
    noise->update(*this);
  }
  noise->evalModel(*this);
  HIPCHK(hipEventRecord(ev_join, side));
  HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
  launch_advance_iter(d_iter.as<int>(), stream);
}

void Model::launch_pending_prior() {
  if (!prior_pending) return;
  const int last = nmodes - 1;
  HIPCHK(hipEventRecord(ev_start, stream));
  HIPCHK(hipStreamWaitEvent(side, ev_start, 0));
  {
    OnSideStream guard(*this);
    priors[last]->update_prior(*this, last);
  }
  launch_advance_iter(d_iter_prior.as<int>(), side);
  HIPCHK(hipEventRecord(ev_join, side));
  HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
  prior_pending = false;
}

void Model::iterate(int n) {
  if (!initialized) throw std::runtime_error("model not initialized");
  bool graphable = use_graph;
  for (auto& p : priors) {
    const auto* mp = dynamic_cast<const MacauPrior*>(p.get());
    if (mp && !mp->capturable()) graphable = false;
  }
  for (int i = 0; i < n; i++) {
    if (graphable) {
      // each graph defers its last prior update to the next launch; the
      // update deferred by the last one is enqueued before iterate returns
      const int v = prior_pending ? 1 : 0;
      if (!graph_exec[v]) {
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        step_launch(true, v == 1);
        HIPCHK(hipStreamEndCapture(stream, &graph[v]));
        HIPCHK(hipGraphInstantiate(&graph_exec[v], graph[v], nullptr, nullptr, 0));
      }
      HIPCHK(hipGraphLaunch(graph_exec[v], stream));
      prior_pending = true;
    } else {
      step_launch();
    }
    host_iter++;
  }
  launch_pending_prior();
}

std::vector<double> Model::profile(int n) {
  if (!initialized) throw std::runtime_error("model not initialized");
  const int nph = 2 * nmodes + 2;
  std::vector<hipEvent_t> ev((size_t)n * (nph + 1));
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int i = 0; i < n; i++) {
    hipEvent_t* E = &ev[(size_t)i * (nph + 1)];
    int q = 0;
    HIPCHK(hipEventRecord(E[q++], stream));
    for (int m = 0; m < nmodes; m++) {
      noise->augment(*this, m);
      priors[m]->sample_latents(*this, m);
      HIPCHK(hipEventRecord(E[q++], stream));
    }
    for (int m = 0; m < nmodes; m++) {
      priors[m]->update_prior(*this, m);
      if (m == nmodes - 1) launch_advance_iter(d_iter_prior.as<int>(), stream);
      HIPCHK(hipEventRecord(E[q++], stream));
    }
    noise->update(*this);
    HIPCHK(hipEventRecord(E[q++], stream));
    noise->evalModel(*this);
    launch_advance_iter(d_iter.as<int>(), stream);
    HIPCHK(hipEventRecord(E[q++], stream));
    host_iter++;
  }
  HIPCHK(hipStreamSynchronize(stream));
  std::vector<double> ms(nph, 0.0);
  for (int i = 0; i < n; i++) {
    hipEvent_t* E = &ev[(size_t)i * (nph + 1)];
    for (int p = 0; p < nph; p++) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, E[p], E[p + 1]));
      ms[p] += t / n;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return ms;
}

void Model::check_fail() {
  const int f = download_scalar<int>(d_fail, stream);
  if (f) {
    upload_scalar<int>(d_fail, 0, stream);
    // chol_decomp's message for the F^T F branch (chol.cpp:21-28), the per-row
    // sampler's otherwise (latentprior.cpp:375)
    if (f == kFailDenseChol) throw std::runtime_error("c++ error: Cholesky decomp failed");
    // the reference's CondNormalWishart does not check its LLTs (its draw
    // would carry NaNs on); here the chain stops with this message
    if (f == kFailNormalWishart)
      throw std::runtime_error("CondNormalWishart: the scale or precision matrix is not positive definite");
    throw std::runtime_error("Cholesky Decomposition failed!");
  }
}

double Model::host_eval_out(int which) {
  double v[2];
  HIPCHK(hipMemcpyAsync(v, d_eval_out.as<void>(), sizeof(v), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return v[which];
}

double Model::factor_norm(int mode) {
  launch_sumsq_f32(d_U[mode].as<float>(), (size_t)KP * dims[mode], d_norm_partial.as<double>(), 256, stream);
  std::vector<double> h(256);
  HIPCHK(hipMemcpyAsync(h.data(), d_norm_partial.as<void>(), 256 * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  double s = 0.0;
  for (double x : h) s += x;
  return std::sqrt(s);
}

double Model::noise_precision() { return download_scalar<double>(d_alpha, stream); }

void Model::download_latents(int mode, std::vector<double>& out) {
  std::vector<float> h((size_t)KP * dims[mode]);
  HIPCHK(hipMemcpyAsync(h.data(), d_U[mode].as<void>(), h.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  out.resize((size_t)num_latent * dims[mode]);
  for (int n = 0; n < dims[mode]; n++)
    for (int k = 0; k < num_latent; k++) out[(size_t)n * num_latent + k] = h[(size_t)n * KP + k];
}

static volatile sig_atomic_t keep_running = 1;
constexpr int kFailPollIters = 10;
static void int_handler(int) {
  keep_running = 0;
  printf("[Received Ctrl-C. Stopping after finishing the current iteration.]\n");
}

void Model::run() {
  init();
  if (verbose) {
    printf("%s\n", noise->getInitStatus().c_str());
    printf("Sampling\n");
  }
  if (save_model) saveGlobalParams();
  keep_running = 1;
  signal(SIGINT, int_handler);
  const double start = tick();
  for (int i = 0; i < burnin + nsamples; i++) {
    if (!keep_running) {
      keep_running = 1;
      break;
    }
    if (verbose && i == burnin) printf(" ====== Burn-in complete, averaging samples ====== \n");
    const double starti = tick();
    iterate(1);
    const bool need_host = verbose || (save_model && i >= burnin);
    if (need_host) {
      sync();
      check_fail();
      noise->rmse_test_onesample = host_eval_out(0);
      noise->rmse_test = host_eval_out(1);
      noise->sync_host(*this);
    }
    if (save_model && i >= burnin) saveModel(i - burnin + 1);
    if (verbose) printStatus(i, tick() - starti);
    // a failed factorisation stops the chain within kFailPollIters iterations
    // (every iteration when the host reads the state anyway)
    if (!need_host && (i + 1) % kFailPollIters == 0) {
      sync();
      check_fail();
    }
  }
  (void)start;
  sync();
  check_fail();
  noise->rmse_test_onesample = host_eval_out(0);
  noise->rmse_test = host_eval_out(1);
  noise->sync_host(*this);
  rmse_test = noise->getEvalMetric();
}

void Model::printStatus(int i, double elapsedi) {  // macau.cpp:151-172
  printf("Iter %3d: %s  ", i, noise->getEvalString().c_str());
  printf("U:[");
  for (int p = 0; p < nmodes; p++) {
    printf("%1.2e", factor_norm(p));
    if (p + 1 < nmodes) printf(", ");
  }
  printf("]");
  printf("\tSide:[");
  for (int p = 0; p < nmodes; p++) {
    printf("%1.2e", priors[p]->getLinkNorm(*this));
    if (p + 1 < nmodes) printf(", ");
  }
  printf("]");
  printf("%s [took %0.1fs]\n", noise->getStatus().c_str(), elapsedi);
  fflush(stdout);
}

std::vector<double> Model::test_to_host(const DevBuf& d) {
  std::vector<double> dev(ntest), out(ntest);
  if (ntest) {
    HIPCHK(hipMemcpyAsync(dev.data(), d.as<void>(), ntest * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  for (long long t = 0; t < ntest; t++) out[test_perm[t]] = dev[t];
  return out;
}

std::vector<double> Model::getPredictions() { return test_to_host(d_pred); }

std::vector<double> Model::getStds() {  // macau.cpp:175-188
  if (nsamples <= 1) return std::vector<double>(ntest, NAN);
  std::vector<double> s = test_to_host(d_pred_var);
  const double inorm = 1.0 / (nsamples - 1);
  for (auto& x : s) x = std::sqrt(x * inorm);
  return s;
}

std::vector<double> Model::getTestData() {  // sparsetensor.cpp:63-76, 178-191
  std::vector<double> out((size_t)ntest * (nmodes + 1));
  for (long long t = 0; t < ntest; t++) {
    for (int m = 0; m < nmodes; m++) out[(size_t)m * ntest + t] = test_idx[(size_t)t * nmodes + m];
    out[(size_t)nmodes * ntest + t] = test_vals[t];
  }
  return out;
}

void Model::saveModel(int isample) {  // macau.cpp:191-198
  const std::string fprefix = save_prefix + "-sample" + std::to_string(isample) + "-";
  for (int i = 0; i < nmodes; i++) {
    std::vector<double> U;
    download_latents(i, U);
    write_csv(fprefix + "U" + std::to_string(i + 1) + "-latents.csv", U, num_latent, dims[i]);
    priors[i]->saveModel(*this, i, fprefix + "U" + std::to_string(i + 1));
  }
}

void Model::saveGlobalParams() {  // macau.cpp:201-205
  write_csv(save_prefix + "-meanvalue.csv", std::vector<double>{mean_value}, 1, 1);
}

// ------------------------------------------------------------------
// priors
// ------------------------------------------------------------------
void ILatentPrior::init(Model& m, int mode) {
  const int K = m.num_latent, KP = m.KP;
  // BPMFPrior::init (latentprior.cpp:84-99): mu = 0, Lambda = 10 I, WI = I, mu0 = 0
  std::vector<double> muh(K, 0.0), Lh((size_t)K * K, 0.0), WIh((size_t)K * K, 0.0);
  std::vector<float> Lf((size_t)KP * KP, 0.f), Lmuf(KP, 0.f);
  for (int i = 0; i < K; i++) {
    Lh[(size_t)i * K + i] = 10.0;
    WIh[(size_t)i * K + i] = 1.0;
  }
  for (int i = 0; i < KP; i++) Lf[(size_t)i * KP + i] = i < K ? 10.f : 1.f;
  upload(mu, muh, m.stream);
  upload(mu0, muh, m.stream);
  upload(Lambda, Lh, m.stream);
  upload(WI, WIh, m.stream);
  upload(Lambda_f, Lf, m.stream);
  upload(Lmu_f, Lmuf, m.stream);
  stats_nblk = hyper_stats_blocks(m.dims[mode]);
  stats.alloc((size_t)stats_nblk * (K + K * K) * sizeof(double));
  if (nw_workspace_bytes(K)) nw_ws.alloc(nw_workspace_bytes(K));
  HIPCHK(hipStreamSynchronize(m.stream));
}

void ILatentPrior::sample_latents(Model& m, int mode) {
  const SampleArgs a = m.sample_args(mode, *this);
  if (m.sampler == SAMPLER_BATCHED) {
    SegStreams seg;
    seg.stream = m.seg_stream;
    seg.fork = m.ev_seg_fork[mode];
    seg.join = m.ev_seg_join[mode];
    launch_sample_batched(a, m.plans[mode], m.stream, seg);
  }
  else launch_sample_plain(a, m.stream);
}

void ILatentPrior::saveModel(Model& m, int, const std::string& prefix) {  // latentprior.cpp:307-309
  std::vector<double> h(m.num_latent);
  HIPCHK(hipMemcpyAsync(h.data(), mu.as<void>(), h.size() * sizeof(double), hipMemcpyDeviceToHost, m.stream));
  HIPCHK(hipStreamSynchronize(m.stream));
  write_csv(prefix + "-latentmean.csv", h, m.num_latent, 1);
}

static NWArgs nw_args(Model& m, ILatentPrior& p, int mode, double nu) {
  NWArgs a;
  a.K = m.num_latent;
  a.KP = m.KP;
  a.partial = p.stats.as<double>();
  a.nblk = 1;  // launch_hyper_stats leaves the totals in row 0
  a.N = m.dims[mode];
  a.kappa = 2.0;  // b0 (latentprior.cpp:97)
  a.mu0 = p.mu0.as<double>();
  a.WI = p.WI.as<double>();
  a.nu = nu;
  a.seed = m.seed;
  a.iter = m.d_iter_prior.as<int>();
  a.mode = mode;
  a.ws = p.nw_ws.as<double>();
  a.mu = p.mu.as<double>();
  a.Lambda = p.Lambda.as<double>();
  a.Lambda_f = p.Lambda_f.as<float>();
  a.Lmu_f = p.Lmu_f.as<float>();
  a.fail = m.d_fail.as<int>();
  return a;
}

void BPMFPrior::update_prior(Model& m, int mode) {  // latentprior.cpp:79-81
  launch_hyper_stats_f32(m.d_U[mode].as<float>(), nullptr, m.KP, m.dims[mode], m.num_latent, stats.as<double>(),
                         stats_nblk, m.stream);
  launch_nw_draw(nw_args(m, *this, mode, (double)m.num_latent), m.stream);  // df = K
}

// ---- MacauPrior ----
static void build_csr(int nrow, const std::vector<int>& r, const std::vector<int>& c, const std::vector<double>& v,
                      std::vector<int>& rp, std::vector<int>& cols, std::vector<double>& vals) {
  // new_csr (libfastsparse): grouped by row, input order kept within a row
  rp.assign(nrow + 1, 0);
  for (int x : r) rp[x + 1]++;
  for (int i = 0; i < nrow; i++) rp[i + 1] += rp[i];
  std::vector<int> pos(rp.begin(), rp.end() - 1);
  cols.resize(r.size());
  vals.resize(r.size());
  for (size_t e = 0; e < r.size(); e++) {
    const int d = pos[r[e]]++;
    cols[d] = c[e];
    vals[d] = v.empty() ? 1.0 : v[e];
  }
}

MacauPrior::MacauPrior(int nr, int nc, const std::vector<int>& rows, const std::vector<int>& cols,
                       const std::vector<double>& vals, bool bin, double lb, double t)
    : nrow(nr), nfeat(nc), lambda_beta0(lb), tol(t), binary(bin) {
  for (size_t e = 0; e < rows.size(); e++)
    if (rows[e] < 0 || rows[e] >= nr || cols[e] < 0 || cols[e] >= nc)
      throw std::runtime_error("side information index out of range");
  build_csr(nr, rows, cols, vals, F_rowptr, F_cols, F_vals);
  build_csr(nc, cols, rows, vals, Ft_rowptr, Ft_cols, Ft_vals);
}

CsrDev MacauPrior::F() const {
  CsrDev c = F_plan;
  c.nrow = nrow;
  c.ncol = nfeat;
  c.row_ptr = dF_rowptr.as<int>();
  c.cols = dF_cols.as<int>();
  c.vals = binary ? nullptr : dF_vals.as<double>();
  return c;
}
CsrDev MacauPrior::Ft() const {
  CsrDev c = Ft_plan;
  c.nrow = nfeat;
  c.ncol = nrow;
  c.row_ptr = dFt_rowptr.as<int>();
  c.cols = dFt_cols.as<int>();
  c.vals = binary ? nullptr : dFt_vals.as<double>();
  return c;
}

// nnz-balanced SpMM work list of a CSR matrix (see CsrDev): rows longer
// than kSpmmSegNnz are cut into segments whose partial sums are added in
// segment order; those segments come first so the long work starts early
// Rows listed in dense_rows (ascending) are left to the dense-block kernel:
// they get no segment and are appended to the split rows with nkc slots
// each, after all other slots (the dense kernel writes them).
// Rows flagged in omit get no work at all.
static CsrDev spmm_plan(const std::vector<int>& rp, int nrmax, std::vector<DevBuf>& store, hipStream_t s,
                        const std::vector<int>* dense_rows = nullptr, int nkc = 0,
                        const std::vector<char>* omit = nullptr) {
  const int nrow = (int)rp.size() - 1;
  std::vector<char> is_dense(nrow, 0);
  if (dense_rows)
    for (int r : *dense_rows) is_dense[r] = 1;
  if (omit)
    for (int r = 0; r < nrow; r++)
      if ((*omit)[r]) is_dense[r] = 2;
  std::vector<int4> seg;
  std::vector<int> mrow, mslot0;
  int slots = 0;
  for (int r = 0; r < nrow; r++) {
    if (is_dense[r] || rp[r + 1] - rp[r] <= kSpmmSegNnz) continue;
    mrow.push_back(r);
    mslot0.push_back(slots);
    for (int b = rp[r]; b < rp[r + 1]; b += kSpmmSegNnz)
      seg.push_back(make_int4(r, b, std::min(b + kSpmmSegNnz, rp[r + 1]), slots++));
  }
  const int dslot0 = slots;
  if (dense_rows)
    for (int r : *dense_rows) {
      mrow.push_back(r);
      mslot0.push_back(slots);
      slots += nkc;
    }
  mslot0.push_back(slots);
  const int nsplit = (int)seg.size();
  // whole rows by decreasing length (ties by row): the long ones form a
  // prefix after the split rows' segments
  std::vector<int> whole;
  for (int r = 0; r < nrow; r++)
    if (!is_dense[r] && rp[r + 1] - rp[r] <= kSpmmSegNnz) whole.push_back(r);
  std::stable_sort(whole.begin(), whole.end(),
                   [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
  int nlong = nsplit, nempty = 0;
  for (int r : whole) {
    seg.push_back(make_int4(r, rp[r], rp[r + 1], -1));
    if (rp[r + 1] - rp[r] >= kSpmmLongNnz) nlong++;
    if (rp[r + 1] == rp[r]) nempty++;
  }
  store.clear();
  store.resize(4);
  upload(store[0], seg, s);
  upload(store[1], mrow.empty() ? std::vector<int>{0} : mrow, s);
  upload(store[2], mslot0, s);
  store[3].alloc((size_t)std::max(slots, 1) * nrmax * sizeof(double));
  CsrDev c;
  c.nseg = (int)seg.size();
  c.nlong = nlong;
  c.nempty = nempty;
  c.seg = store[0].as<int4>();
  c.nmulti = (int)mrow.size();
  c.multi_row = store[1].as<int>();
  c.multi_slot0 = store[2].as<int>();
  c.partial = store[3].as<double>();
  c.dslot0 = dslot0;
  c.nkc = nkc;
  return c;
}

// The dense column block of a binary F (or one whose values are all 1): columns with a nonzero in at least
// every 64th row and more than kSpmmSegNnz nonzeros, the densest
// kDenseMaxCols of them, when there are at least 16; none when one of their
// entries is repeated (a bit cannot count it twice).
void MacauPrior::build_dense(int nrmax, hipStream_t s) {
  Fd_plan = CsrDev();
  Ftd_plan = CsrDev();
  dense_bufs.clear();
  Ftd_plan_bufs.clear();
  // values all 1 multiply like a binary F
  const bool ones = binary || std::all_of(F_vals.begin(), F_vals.end(), [](double v) { return v == 1.0; });
  if (!ones || nrmax < 32) return;
  std::vector<int> cand;
  for (int j = 0; j < nfeat; j++) {
    const long c = Ft_rowptr[j + 1] - Ft_rowptr[j];
    if (c * 64 >= nrow && c > kSpmmSegNnz) cand.push_back(j);
  }
  if (cand.size() < 16) return;
  std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) {
    return Ft_rowptr[a + 1] - Ft_rowptr[a] > Ft_rowptr[b + 1] - Ft_rowptr[b];
  });
  if ((int)cand.size() > kDenseMaxCols) cand.resize(kDenseMaxCols);
  std::sort(cand.begin(), cand.end());
  const int nd = (int)cand.size(), ndp = (nd + 15) / 16 * 16;
  const int ndw = (ndp + 31) / 32, nrw = (nrow + 31) / 32;
  std::vector<int> pos(nfeat, -1);
  for (int j = 0; j < nd; j++) pos[cand[j]] = j;
  std::vector<int> dcol(ndp, cand[0]);
  std::copy(cand.begin(), cand.end(), dcol.begin());
  std::vector<uint32_t> rbits((size_t)nrow * ndw, 0u), cbits((size_t)ndp * nrw, 0u);
  std::vector<int> rptr(nrow + 1, 0), rcols;
  for (int r = 0; r < nrow; r++) {
    for (int e = F_rowptr[r]; e < F_rowptr[r + 1]; e++) {
      const int j = pos[F_cols[e]];
      if (j < 0) {
        rcols.push_back(F_cols[e]);
        continue;
      }
      uint32_t& w = rbits[(size_t)r * ndw + j / 32];
      if (w & (1u << (j % 32))) return;  // repeated entry
      w |= 1u << (j % 32);
      cbits[(size_t)j * nrw + r / 32] |= 1u << (r % 32);
    }
    rptr[r + 1] = (int)rcols.size();
  }
  std::vector<int> rfix((size_t)nrow * kDenseRem, 0);
  for (int r = 0; r < nrow; r++)
    for (int t = 0; t < kDenseRem && rptr[r] + t < rptr[r + 1]; t++) rfix[(size_t)r * kDenseRem + t] = rcols[rptr[r] + t];
  if (rcols.empty()) rcols.push_back(0);
  dense_bufs.resize(7);
  upload(dense_bufs[5], rfix, s);
  upload(dense_bufs[0], dcol, s);
  upload(dense_bufs[1], rbits, s);
  upload(dense_bufs[2], rptr, s);
  upload(dense_bufs[3], rcols, s);
  upload(dense_bufs[4], cbits, s);
  CsrDev d;
  d.ndp = ndp;
  d.nd = nd;
  d.nfrow = nrow;
  d.ndw = ndw;
  d.nrw = nrw;
  d.dcol = dense_bufs[0].as<int>();
  Fd_plan = d;
  Fd_plan.dense_role = 1;
  Fd_plan.row_bits = dense_bufs[1].as<uint32_t>();
  Fd_plan.rem_ptr = dense_bufs[2].as<int>();
  Fd_plan.rem_cols = dense_bufs[3].as<int>();
  Fd_plan.rem_fix = dense_bufs[5].as<int>();
  const int nkc = (nrow + kDenseRowChunk - 1) / kDenseRowChunk;
  // columns of F without nonzeros, none within kEmptyReach columns either
  // way: no work unit (the CG forms their KP rows as reg * P)
  std::vector<int> nxt(nfeat + 1, nfeat);
  for (int j = nfeat - 1; j >= 0; j--) nxt[j] = Ft_rowptr[j + 1] > Ft_rowptr[j] ? j : nxt[j + 1];
  std::vector<char> omit(nfeat, 0);
  {
    int prev = -(1 << 30);  // the last column with nonzeros before j
    for (int j = 0; j < nfeat; j++) {
      if (Ft_rowptr[j + 1] > Ft_rowptr[j]) {
        prev = j;
        continue;
      }
      omit[j] = (j - prev > kEmptyReach && nxt[j] - j > kEmptyReach) ? 1 : 0;
    }
  }
  upload(dense_bufs[6], nxt, s);
  Ftd_plan = spmm_plan(Ft_rowptr, nrmax, Ftd_plan_bufs, s, &cand, nkc, &omit);
  Ftd_plan.next_nonempty = dense_bufs[6].as<int>();
  Ftd_plan.dense_role = 2;
  Ftd_plan.ndp = ndp;
  Ftd_plan.nd = nd;
  Ftd_plan.nfrow = nrow;
  Ftd_plan.ndw = ndw;
  Ftd_plan.nrw = nrw;
  Ftd_plan.dcol = d.dcol;
  Ftd_plan.col_bits = dense_bufs[4].as<uint32_t>();
}

CsrDev MacauPrior::Fd() const {
  if (!Fd_plan.dense_role) return CsrDev();
  CsrDev c = Fd_plan;
  const CsrDev base = F();
  c.nrow = base.nrow;
  c.ncol = base.ncol;
  c.row_ptr = base.row_ptr;
  c.cols = base.cols;
  c.vals = nullptr;
  return c;
}

CsrDev MacauPrior::Ftd() const {
  if (!Ftd_plan.dense_role) return CsrDev();
  CsrDev c = Ftd_plan;
  const CsrDev base = Ft();
  c.nrow = base.nrow;
  c.ncol = base.ncol;
  c.row_ptr = base.row_ptr;
  c.cols = base.cols;
  c.vals = nullptr;
  return c;
}

void MacauPrior::upload_side(int nrmax, hipStream_t s) {
  upload(dF_rowptr, F_rowptr, s);
  upload(dF_cols, F_cols, s);
  upload(dF_vals, F_vals, s);
  upload(dFt_rowptr, Ft_rowptr, s);
  upload(dFt_cols, Ft_cols, s);
  upload(dFt_vals, Ft_vals, s);
  F_plan = spmm_plan(F_rowptr, nrmax, F_plan_bufs, s);
  Ft_plan = spmm_plan(Ft_rowptr, nrmax, Ft_plan_bufs, s);
  build_dense(nrmax, s);
}

void MacauPrior::init(Model& m, int mode) {
  ILatentPrior::init(m, mode);
  if (nrow != m.dims[mode])
    throw std::runtime_error("side information rows must equal the size of its data mode");
  const int K = m.num_latent, KP = m.KP, N = nrow;
  upload_side(std::max(K, 40), m.stream);
  const size_t kf = (size_t)K * nfeat;
  beta.alloc(kf * sizeof(double));
  HIPCHK(hipMemsetAsync(beta.as<void>(), 0, beta.bytes(), m.stream));
  Uhat.alloc((size_t)KP * N * sizeof(float));
  HIPCHK(hipMemsetAsync(Uhat.as<void>(), 0, Uhat.bytes(), m.stream));
  Lmu_rows.alloc((size_t)KP * N * sizeof(float));
  HIPCHK(hipMemsetAsync(Lmu_rows.as<void>(), 0, Lmu_rows.bytes(), m.stream));  // Lambda (mu + 0), mu = 0
  upload_scalar<double>(lambda_beta_d, lambda_beta0, m.stream);
  Lchol.alloc((size_t)K * K * sizeof(double));
  tmp.alloc((size_t)K * N * sizeof(double));
  Fty.alloc(kf * sizeof(double));
  bbt_nblk = hyper_stats_blocks(nfeat);
  bbt.alloc((size_t)bbt_nblk * (K + K * K) * sizeof(double));
  partial_sumsq.alloc(256 * sizeof(double));
  // block CG buffers sized for the largest block (<= 40 right-hand sides)
  cg.alloc(std::min(K, 40), nfeat, N);
  if (use_ftf) {  // FtF = At_mul_A(F) once (latentprior.cpp:199-203)
    const size_t nn = (size_t)nfeat * nfeat * sizeof(double);
    FtF.alloc(nn);
    Wchol.alloc(nn);
    launch_ftf(F(), Ft(), FtF.as<double>(), m.stream);
    ftf_solver = kFtfCholesky;
    const char* env = std::getenv("MACAU_AMD_FTF_SOLVER");
    if (env && std::strcmp(env, "cholesky") == 0) {
      ftf_note = "MACAU_AMD_FTF_SOLVER=cholesky";
    } else {
      ftf_d.alloc((size_t)nfeat * sizeof(double));
      if (spectral_factor(FtF.as<double>(), nfeat, ftf_d.as<double>(), Wchol.as<double>(), m.stream, &ftf_note)) {
        ftf_solver = kFtfSpectral;
        ftf_plan = spectral_plan(nfeat, K);
        ftf_y.alloc(kf * sizeof(double));
        ftf_part.alloc((size_t)ftf_plan.S * kf * sizeof(double));
      } else {
        // the eigensolver may have overwritten F^T F: form it again
        launch_ftf(F(), Ft(), FtF.as<double>(), m.stream);
      }
    }
  }
  HIPCHK(hipStreamSynchronize(m.stream));
}

void CgBuffers::alloc(int nr, int nfeat, int nsamples) {
  nrmax = nr;
  X.alloc((size_t)nr * nfeat * sizeof(double));
  R.alloc((size_t)nr * nfeat * sizeof(double));
  P.alloc((size_t)nr * nfeat * sizeof(double));
  KP.alloc((size_t)nr * nfeat * sizeof(double));
  B.alloc((size_t)nr * nfeat * sizeof(double));
  inner.alloc((size_t)nr * std::max(nsamples, 1) * sizeof(double));
  partial.alloc((size_t)dots_blocks(nfeat) * nr * nr * sizeof(double));
  small.alloc((size_t)(5 * nr * nr + nr + 4) * sizeof(double));
}

CgBuffers::~CgBuffers() {
  for (auto& c : chunks)
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
}

int blockcg_device(const CsrDev& F0, const CsrDev& Ft0, const double* reg, const double* B, int ld, int r0, int nr,
                   double tol, double* X, CgBuffers& b, hipStream_t s, bool use_graph, const CsrDev* Fd,
                   const CsrDev* Ftd) {
  // 32 right-hand sides: the dense-block products when F has them
  const bool dense = nr == 32 && Fd && Ftd && Fd->dense_role == 1 && Ftd->dense_role == 2;
  const CsrDev& F = dense ? *Fd : F0;
  const CsrDev& Ft = dense ? *Ftd : Ft0;
  if (nr > b.nrmax || nr > 40) throw std::runtime_error("BlockCG only available for up to 40 RHSs.");
  const int nfeat = F.ncol;
  CgState c;
  c.nr = nr;
  c.n = nfeat;
  c.X = (nr == ld) ? X : b.X.as<double>();
  c.R = b.R.as<double>();
  c.P = b.P.as<double>();
  c.KPm = b.KP.as<double>();
  c.inner = b.inner.as<double>();
  c.partial = b.partial.as<double>();
  c.nblk = dots_blocks(nfeat);
  c.small = b.small.as<double>();
  if (dense) {
    c.nxt = Ft.next_nonempty;
    c.reg = reg;
  }
  const double* Bb = B;
  if (nr != ld) {
    launch_rows_copy(B, ld, b.B.as<double>(), nr, r0, 0, nr, nfeat, s);
    Bb = b.B.as<double>();
  }
  launch_cg_init(c, Bb, s);
  const double tolsq = tol * tol;
  double flags[3] = {0, 0, 0};
  const double* dflags = c.small + 5 * nr * nr + nr;
  // a chunk of iterations per host round trip: the first one iteration
  // shorter than the previous solve needed (8 without one), then 2 at a
  // time; kernels of iterations after convergence return immediately
  // (device-side done flag)
  auto launch_chunk = [&](hipStream_t st, int len) {
    for (int q = 0; q < len; q++) {
      launch_spmm(F, c.P, nr, c.inner, nullptr, nullptr, st, dflags);
      launch_spmm(Ft, c.inner, nr, c.KPm, reg, c.P, st, dflags);
      launch_cg_step_a(c, st);
      launch_cg_update_xr(c, st);
      launch_cg_step_b(c, tolsq, st);
      launch_cg_update_p(c, st);
    }
  };
  std::vector<uintptr_t> key;
  if (use_graph) {
    double tolsq_bits = tolsq;
    uintptr_t tb = 0;
    std::memcpy(&tb, &tolsq_bits, sizeof(tb));
    key = {(uintptr_t)reg, (uintptr_t)c.X, (uintptr_t)c.small, (uintptr_t)nr, tb};
    for (const CsrDev* A : {&F, &Ft}) {
      const uintptr_t f[] = {(uintptr_t)A->nrow,       (uintptr_t)A->ncol,      (uintptr_t)A->row_ptr,
                             (uintptr_t)A->cols,       (uintptr_t)A->vals,      (uintptr_t)A->nseg,
                             (uintptr_t)A->seg,        (uintptr_t)A->nmulti,    (uintptr_t)A->multi_row,
                             (uintptr_t)A->multi_slot0, (uintptr_t)A->partial,  (uintptr_t)A->dense_role,
                             (uintptr_t)A->dcol,       (uintptr_t)A->row_bits,  (uintptr_t)A->col_bits};
      key.insert(key.end(), std::begin(f), std::end(f));
    }
  }
  // the chunk's graph for this configuration and length, captured once
  auto chunk_exec = [&](int len) -> hipGraphExec_t {
    std::vector<uintptr_t> k = key;
    k.push_back((uintptr_t)len);
    for (auto& ch : b.chunks)
      if (ch.key == k) return ch.exec;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_chunk(s, len);
    HIPCHK(hipStreamEndCapture(s, &g));
    HIPCHK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    b.chunks.push_back({k, exec});
    return exec;
  };
  int len = b.last_iters >= 8 ? b.last_iters - 1 : 8;
  for (int it = 0; it < 100000; it += len, len = 2) {
    if (use_graph)
      HIPCHK(hipGraphLaunch(chunk_exec(len), s));
    else
      launch_chunk(s, len);
    HIPCHK(hipMemcpyAsync(flags, dflags, sizeof(flags), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flags[0] != 0.0) break;
  }
  b.last_iters = (int)flags[1];
  if (flags[2] != 0.0) throw std::runtime_error("block CG: Cholesky of a small system failed");
  launch_cg_finish(c, s);
  if (nr != ld) launch_rows_copy(c.X, nr, X, ld, 0, r0, nr, nfeat, s);
  return (int)flags[1];
}

void MacauPrior::update_prior(Model& m, int mode) {  // latentprior.cpp:254-264
  const int K = m.num_latent, KP = m.KP, N = nrow;
  hipStream_t s = m.stream;
  // residual statistics of U - Uhat, and beta beta^T of the current beta
  launch_hyper_stats_f32(m.d_U[mode].as<float>(), Uhat.as<float>(), KP, N, K, stats.as<double>(), stats_nblk, s);
  launch_hyper_stats_f64(beta.as<double>(), K, nfeat, K, bbt.as<double>(), bbt_nblk, s);
  NWArgs a = nw_args(m, *this, mode, (double)K + (double)nfeat);  // df + beta.cols()
  a.BBt_partial = bbt.as<double>();
  a.bbt_nblk = 1;
  a.lambda_beta = lambda_beta_d.as<double>();
  a.Lchol = Lchol.as<double>();
  launch_nw_draw(a, s);
  // sample_beta (latentprior.cpp:272-293)
  BetaArgs b;
  b.K = K;
  b.KP = KP;
  b.N = N;
  b.nfeat = nfeat;
  b.U = m.d_U[mode].as<float>();
  b.mu = mu.as<double>();
  b.Lchol = Lchol.as<double>();
  b.lambda_beta = lambda_beta_d.as<double>();
  b.seed = m.seed;
  b.iter = m.d_iter_prior.as<int>();
  b.mode = mode;
  launch_beta_tmp(b, tmp.as<double>(), s);
  launch_spmm(Ft(), tmp.as<double>(), K, Fty.as<double>(), nullptr, nullptr, s);
  launch_beta_noise_add(b, Fty.as<double>(), s);
  if (use_ftf && ftf_solver == kFtfSpectral) {
    // beta = Ft_y (FtF + lambda_beta I)^{-1} (latentprior.cpp:280-288)
    // through the eigendecomposition of FtF made at init
    launch_spectral_solve(FtF.as<double>(), Wchol.as<double>(), ftf_d.as<double>(), nfeat, lambda_beta_d.as<double>(),
                          Fty.as<double>(), ftf_y.as<double>(), ftf_part.as<double>(), ftf_plan, beta.as<double>(), K,
                          m.d_fail.as<int>(), s);
    last_cg_iters = 0;
  } else if (use_ftf) {
    // K = FtF + lambda_beta I; chol_decomp(K); chol_solve_t(K, Ft_y);
    // beta = Ft_y (latentprior.cpp:280-288)
    HIPCHK(hipMemcpyAsync(beta.as<void>(), Fty.as<void>(), (size_t)K * nfeat * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
    launch_ftf_solve(FtF.as<double>(), nfeat, lambda_beta_d.as<double>(), Wchol.as<double>(), beta.as<double>(), K,
                     m.d_fail.as<int>(), s);
    last_cg_iters = 0;
  } else {
    // solve_blockcg(beta, F, lambda_beta, Ft_y, tol, 32, 8) (linop.h:206-226)
    const int blocksize = 32, excess = 8;
    const CsrDev fd = Fd(), ftd = Ftd();
    if (K <= excess + blocksize) {
      last_cg_iters = blockcg_device(F(), Ft(), lambda_beta_d.as<double>(), Fty.as<double>(), K, 0, K, tol,
                                     beta.as<double>(), cg, s, m.use_graph, &fd, &ftd);
    } else {
      for (int i = 0; i < K; i += blocksize) {
        int nr = blocksize;
        if (i + blocksize + excess >= K) nr = K - i;
        last_cg_iters = blockcg_device(F(), Ft(), lambda_beta_d.as<double>(), Fty.as<double>(), K, i, nr, tol,
                                       beta.as<double>(), cg, s, m.use_graph, &fd, &ftd);
      }
    }
  }
  // compute_uhat and the per-row prior term for the next half-sweep
  launch_uhat_finish(F(), beta.as<double>(), K, KP, mu.as<double>(), Lambda.as<double>(), Uhat.as<float>(),
                     Lmu_rows.as<float>(), tmp.as<double>(), s);
  // lambda_beta = sample_lambda_beta(beta, Lambda, nu0, mu0)
  launch_hyper_stats_f64(beta.as<double>(), K, nfeat, K, bbt.as<double>(), bbt_nblk, s);
  launch_lambda_beta(K, (double)nfeat, bbt.as<double>(), 1, Lambda.as<double>(), lambda_beta_d.as<double>(),
                     m.seed, m.d_iter_prior.as<int>(), mode, s);
}

double MacauPrior::getLinkNorm(Model& m) {
  launch_sumsq_f64(beta.as<double>(), (size_t)m.num_latent * nfeat, partial_sumsq.as<double>(), 256, m.stream);
  std::vector<double> h(256);
  HIPCHK(hipMemcpyAsync(h.data(), partial_sumsq.as<void>(), 256 * sizeof(double), hipMemcpyDeviceToHost, m.stream));
  HIPCHK(hipStreamSynchronize(m.stream));
  double s = 0.0;
  for (double x : h) s += x;
  return std::sqrt(s);
}

double MacauPrior::getLinkLambda(Model& m) { return download_scalar<double>(lambda_beta_d, m.stream); }

void MacauPrior::saveModel(Model& m, int mode, const std::string& prefix) {  // latentprior.cpp:311-315
  ILatentPrior::saveModel(m, mode, prefix);
  std::vector<double> h((size_t)m.num_latent * nfeat);
  HIPCHK(hipMemcpyAsync(h.data(), beta.as<void>(), h.size() * sizeof(double), hipMemcpyDeviceToHost, m.stream));
  HIPCHK(hipStreamSynchronize(m.stream));
  write_csv(prefix + "-link.csv", h, m.num_latent, nfeat);
}

// ---- MacauOnePrior ----
MacauOnePrior::MacauOnePrior(int nr, int nc, const std::vector<int>& rows, const std::vector<int>& cols,
                             const std::vector<double>& vals, bool bin, double lb)
    : MacauPrior(nr, nc, rows, cols, vals, bin, lb, 1e-6) {
  // col_square_sum (linop.cpp:374-401): entries counted (binary) or squared
  colsq.assign(nc, 0.0);
  for (int f = 0; f < nc; f++)
    for (int e = Ft_rowptr[f]; e < Ft_rowptr[f + 1]; e++) colsq[f] += Ft_vals[e] * Ft_vals[e];
  // F^T with the repeated entries of each row merged (values added), so the
  // sweep's scatter has distinct rows per column; At_mul_Bt and
  // add_Acol_mul_bt (linop.h:121-178) give the same sums
  Fu_rowptr.assign(nc + 1, 0);
  Fu_cols.clear();
  Fu_vals.clear();
  std::vector<std::pair<int, double>> ent;
  for (int f = 0; f < nc; f++) {
    ent.clear();
    for (int e = Ft_rowptr[f]; e < Ft_rowptr[f + 1]; e++) ent.emplace_back(Ft_cols[e], Ft_vals[e]);
    std::stable_sort(ent.begin(), ent.end(), [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
      return x.first < y.first;
    });
    for (size_t q = 0; q < ent.size(); q++) {
      if (q > 0 && ent[q].first == ent[q - 1].first) {
        Fu_vals.back() += ent[q].second;
      } else {
        Fu_cols.push_back(ent[q].first);
        Fu_vals.push_back(ent[q].second);
      }
    }
    Fu_rowptr[f + 1] = (int)Fu_cols.size();
  }
  // the sweep's batches (greedy in feature order; any order-preserving
  // grouping of commuting updates gives the sequential sweep's result):
  // features with empty columns are left to the parallel pre-pass, columns
  // longer than 16 entries run alone on the whole wave, and up to four
  // consecutive shorter columns whose rows are pairwise disjoint share a step
  std::vector<int> stamp(nr, -1), cur;
  int nbat = 0;
  auto flush = [&]() {
    if (cur.empty()) return;
    for (int q = 0; q < 4; q++) {
      if (q < (int)cur.size())
        batch_meta.push_back(make_int4(cur[q], Fu_rowptr[cur[q]], Fu_rowptr[cur[q] + 1], 0));
      else
        batch_meta.push_back(make_int4(-1, 0, 0, 0));
    }
    cur.clear();
    nbat++;
  };
  for (int f = 0; f < nc; f++) {
    const int len = Fu_rowptr[f + 1] - Fu_rowptr[f];
    if (len == 0) continue;
    if (len > 16) {
      flush();
      for (int q = 0; q < 4; q++) batch_meta.push_back(make_int4(f, Fu_rowptr[f], Fu_rowptr[f + 1], 1));
      nbat++;
      continue;
    }
    bool conflict = cur.size() == 4;
    for (int e = Fu_rowptr[f]; e < Fu_rowptr[f + 1] && !conflict; e++) conflict = stamp[Fu_cols[e]] == nbat;
    if (conflict) flush();
    cur.push_back(f);
    for (int e = Fu_rowptr[f]; e < Fu_rowptr[f + 1]; e++) stamp[Fu_cols[e]] = nbat;
  }
  flush();
}

void MacauOnePrior::init(Model& m, int mode) {  // macauoneprior.cpp:16-38
  ILatentPrior::init(m, mode);  // mu = 0, Lambda = 10 I (lambda = 10)
  if (nrow != m.dims[mode])
    throw std::runtime_error("side information rows must equal the size of its data mode");
  const int K = m.num_latent, KP = m.KP, N = nrow;
  upload_side(K, m.stream);
  const size_t kf = (size_t)K * nfeat;
  beta.alloc(std::max<size_t>(kf, 1) * sizeof(double));
  HIPCHK(hipMemsetAsync(beta.as<void>(), 0, beta.bytes(), m.stream));
  Uhat.alloc((size_t)KP * N * sizeof(float));
  HIPCHK(hipMemsetAsync(Uhat.as<void>(), 0, Uhat.bytes(), m.stream));
  Lmu_rows.alloc((size_t)KP * N * sizeof(float));
  HIPCHK(hipMemsetAsync(Lmu_rows.as<void>(), 0, Lmu_rows.bytes(), m.stream));
  uhat64.alloc(std::max<size_t>((size_t)K * N, 1) * sizeof(double));
  HIPCHK(hipMemsetAsync(uhat64.as<void>(), 0, uhat64.bytes(), m.stream));
  upload(lambda_beta_vec, std::vector<double>(K, lambda_beta0), m.stream);  // setLambdaBeta
  upload(dcolsq, colsq.empty() ? std::vector<double>{0.0} : colsq, m.stream);
  upload(dFu_rowptr, Fu_rowptr, m.stream);
  upload(dFu_cols, Fu_cols.empty() ? std::vector<int>{0} : Fu_cols, m.stream);
  upload(dFu_vals, Fu_vals.empty() ? std::vector<double>{0.0} : Fu_vals, m.stream);
  upload(dbatch, batch_meta.empty() ? std::vector<int4>{make_int4(-1, 0, 0, 0)} : batch_meta, m.stream);
  ainv.alloc(std::max<size_t>(kf, 1) * sizeof(double));
  nz.alloc(std::max<size_t>(kf, 1) * sizeof(double));
  Zw.alloc(beta_uni_in_lds(N) ? sizeof(double) : (size_t)K * N * sizeof(double));
  lb_partial.alloc((size_t)lambda_beta_uni_partials(K) * sizeof(double));
  partial_sumsq.alloc(256 * sizeof(double));
  HIPCHK(hipStreamSynchronize(m.stream));
}

void MacauOnePrior::sample_latents(Model& m, int mode) {  // macauoneprior.cpp:42-95, 108-187
  SampleArgs a = m.sample_args(mode, *this);
  a.univariate = 1;
  launch_sample_plain(a, m.stream);
}

void MacauOnePrior::update_prior(Model& m, int mode) {  // macauoneprior.cpp:189-196
  const int K = m.num_latent, KP = m.KP, N = nrow;
  hipStream_t s = m.stream;
  // sample_mu_lambda (:198-213): CondNormalWishart(U - Uhat, 0, 2, I, K), diagonal kept
  launch_hyper_stats_f32(m.d_U[mode].as<float>(), Uhat.as<float>(), KP, N, K, stats.as<double>(), stats_nblk, s);
  launch_nw_draw(nw_args(m, *this, mode, (double)K), s);
  launch_uni_diag(K, KP, Lambda.as<double>(), Lambda_f.as<float>(), Lmu_f.as<float>(), mu.as<double>(), s);
  // sample_beta (:215-252) from the previous Uhat
  BetaUniArgs b;
  b.K = K;
  b.KP = KP;
  b.N = N;
  b.nfeat = nfeat;
  b.U = m.d_U[mode].as<float>();
  b.uhat = uhat64.as<double>();
  b.mu = mu.as<double>();
  b.Lambda = Lambda.as<double>();
  b.lambda_beta = lambda_beta_vec.as<double>();
  b.colsq = dcolsq.as<double>();
  b.Ft_rowptr = dFu_rowptr.as<int>();
  b.Ft_cols = dFu_cols.as<int>();
  b.Ft_vals = dFu_vals.as<double>();
  b.beta = beta.as<double>();
  b.ainv = ainv.as<double>();
  b.nz = nz.as<double>();
  b.Zw = Zw.as<double>();
  b.batch_meta = dbatch.as<int4>();
  b.nbatch = (int)(batch_meta.size() / 4);
  b.seed = m.seed;
  b.iter = m.d_iter_prior.as<int>();
  b.mode = mode;
  launch_beta_uni(b, s);
  // compute_uhat, with the per-row prior term lambda (mu + Uhat_n) for the
  // next half-sweep; the FP64 Uhat stays in uhat64 for the next sweep
  launch_uhat_finish(F(), beta.as<double>(), K, KP, mu.as<double>(), Lambda.as<double>(), Uhat.as<float>(),
                     Lmu_rows.as<float>(), uhat64.as<double>(), s);
  // sample_lambda_beta (:263-289)
  launch_lambda_beta_uni(b, lb_partial.as<double>(), s);
}

double MacauOnePrior::getLinkLambda(Model& m) {
  std::vector<double> h(m.num_latent);
  HIPCHK(hipMemcpyAsync(h.data(), lambda_beta_vec.as<void>(), h.size() * sizeof(double), hipMemcpyDeviceToHost,
                        m.stream));
  HIPCHK(hipStreamSynchronize(m.stream));
  double s = 0.0;
  for (double x : h) s += x;
  return h.empty() ? NAN : s / (double)h.size();
}

// ------------------------------------------------------------------
// noise models
// ------------------------------------------------------------------
std::string INoiseModel::getEvalString() {  // noisemodels.h:46
  return std::string("RMSE: ") + to_string_with_precision(rmse_test, 5) + " (1samp: " +
         to_string_with_precision(rmse_test_onesample, 5) + ")";
}

static EvalArgs eval_args(Model& m) {
  EvalArgs a;
  a.ntest = (int)m.ntest;
  a.nmodes = m.nmodes;
  a.K = m.num_latent;
  a.KP = m.KP;
  a.idx = m.d_test_idx.as<int>();
  a.vals = m.d_test_vals.as<double>();
  for (int q = 0; q < m.nmodes; q++) a.f[q] = m.d_U[q].as<float>();
  a.mean = m.mean_value;
  a.pred = m.d_pred.as<double>();
  a.pred_var = m.d_pred_var.as<double>();
  a.partial = m.d_eval_partial.as<double>();
  a.nblk = m.eval_nblk;
  a.iter = m.d_iter.as<int>();
  a.burnin = m.burnin;
  a.out = m.d_eval_out.as<double>();
  return a;
}

void INoiseModel::evalModel(Model& m) { launch_eval(eval_args(m), m.stream); }

void FixedGaussianNoise::init(Model& m) { upload_scalar<double>(m.d_alpha, alpha, m.stream); }

std::string FixedGaussianNoise::getInitStatus() {
  return std::string("Noise precision: ") + std::to_string(alpha) + " (fixed)";
}

void AdaptiveGaussianNoise::init(Model& m) {  // noisemodels.cpp:14-53
  double se = 0.0;
  for (double v : m.rows[0].vals) se += (v - m.mean_value) * (v - m.mean_value);
  var_total = se / (double)m.nnz;
  if (var_total <= 0.0 || std::isnan(var_total)) var_total = 1.0;
  alpha = (sn_init + 1.0) / var_total;
  alpha_max = (sn_max + 1.0) / var_total;
  upload_scalar<double>(m.d_alpha, alpha, m.stream);
}

void AdaptiveGaussianNoise::update(Model& m) {  // noisemodels.cpp:55-124
  NoiseDrawArgs d;
  d.nnz = (double)m.nnz;
  d.var_total = var_total;
  d.b0 = m.is_tensor ? 0.5 : 0.5 * var_total;
  d.alpha_max = alpha_max;
  d.seed = m.seed;
  d.iter = m.d_iter.as<int>();
  d.alpha = m.d_alpha.as<double>();
  d.sse_out = m.d_sse.as<double>();
  if (m.fused_sse) {  // SSE already reduced per wave by the last half-sweep
    d.partial = m.d_sse_partial.as<double>();
    d.nblk = m.sse_partial_n;
    launch_noise_draw(d, m.stream);
    return;
  }
  NoiseArgs a;
  a.rows0 = m.rdev[0];
  a.U0 = m.d_U[0].as<float>();
  a.others = m.others_of(0);
  a.K = m.num_latent;
  a.KP = m.KP;
  a.nmodes = m.nmodes;
  a.partial = m.d_noise_partial.as<double>();
  a.nblk = m.noise_nblk;
  launch_noise_sse(a, m.stream);
  d.partial = m.d_noise_partial.as<double>();
  d.nblk = m.noise_nblk;
  launch_noise_draw(d, m.stream);
}

std::string AdaptiveGaussianNoise::getInitStatus() {
  return std::string("Noise precision: adaptive (with max precision of ") + std::to_string(alpha_max) + ")";
}

std::string AdaptiveGaussianNoise::getStatus() { return std::string("Prec:") + to_string_with_precision(alpha, 2); }

void AdaptiveGaussianNoise::sync_host(Model& m) { alpha = m.noise_precision(); }

// ------------------------------------------------------------------
// ProbitNoise
// ------------------------------------------------------------------
void ProbitNoise::init(Model& m) {
  if (m.nmodes != 2) {
    // BPMFPrior::sample_latents(ProbitNoise&, TensorData&) throws
    // (latentprior.cpp:112-116); with side information the sampler does
    // nothing and ProbitNoise::evalModel(TensorData) throws
    // (noisemodels.cpp:177-181)
    for (auto& p : m.priors)
      if (!dynamic_cast<MacauPrior*>(p.get())) throw std::runtime_error("Unimplemented: sample_latents");
    throw std::runtime_error("ProbitNoise::evalModel unimplemented.");
  }
  upload_scalar<double>(m.d_alpha, 1.0, m.stream);
  d_y.clear();
  d_row_of.clear();
  d_z.clear();
  for (int q = 0; q < m.nmodes; q++) {
    const HostRows& h = m.rows[q];
    std::vector<int> row_of(h.vals.size());
    for (int n = 0; n < m.dims[q]; n++)
      for (int e = h.row_ptr[n]; e < h.row_ptr[n + 1]; e++) row_of[e] = n;
    d_y.emplace_back();
    d_row_of.emplace_back();
    upload(d_y.back(), h.vals, m.stream);
    upload(d_row_of.back(), row_of, m.stream);
    d_z.emplace_back(std::max<size_t>(h.vals.size(), 1) * sizeof(float));
    m.rdev[q].r = d_z.back().as<float>();  // the samplers read z in place of y - mean
  }
  d_pred1.alloc(std::max<long long>(m.ntest, 1) * sizeof(double));
  HIPCHK(hipStreamSynchronize(m.stream));
}

void ProbitNoise::augment(Model& m, int mode) {
  ProbitZArgs a;
  a.nnz = (long long)m.rows[mode].vals.size();
  a.KP = m.KP;
  a.row_of = d_row_of[mode].as<int>();
  a.idx = m.d_idx[mode].as<int>();
  a.y = d_y[mode].as<double>();
  a.U = m.d_U[mode].as<float>();
  a.V = m.d_U[1 - mode].as<float>();
  a.z = d_z[mode].as<float>();
  a.seed = m.seed;
  a.iter = m.d_iter.as<int>();
  a.mode = mode;
  launch_probit_z(a, m.stream);
}

void ProbitNoise::evalModel(Model& m) {
  EvalArgs a = eval_args(m);
  a.probit = 1;
  a.pred1 = d_pred1.as<double>();
  launch_eval(a, m.stream);
}

std::string ProbitNoise::getEvalString() {  // noisemodels.h:100
  return std::string("AUC: ") + to_string_with_precision(auc_test, 5) + " (1samp: " +
         to_string_with_precision(auc_test_onesample, 5) + ")";
}

void ProbitNoise::sync_host(Model& m) {
  const std::vector<double> pred = m.test_to_host(m.d_pred), pred1 = m.test_to_host(d_pred1);
  auc_test_onesample = auc(pred1, m.test_vals);
  auc_test = auc(pred, m.test_vals);
}

double auc(const std::vector<double>& pred, const std::vector<double>& test) {
  const size_t n = pred.size();
  if (n == 0) return NAN;
  std::vector<size_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](size_t i, size_t j) { return pred[i] < pred[j]; });
  double np = 0.0;
  for (double t : test) np += t;
  const double nn = (double)n - np;
  // x: positives so far, y: negatives so far, along ascending predictions;
  // each step adds (x_{i+1} - x_i) y_{i+1}
  double x = test[perm[0]], y = 1.0 - x, area = 0.0;
  for (size_t i = 1; i < n; i++) {
    const double x1 = x + test[perm[i]], y1 = y + 1.0 - test[perm[i]];
    area += (x1 - x) * y1;
    x = x1;
    y = y1;
  }
  return area / (np * nn);
}

}  // namespace mgpu
