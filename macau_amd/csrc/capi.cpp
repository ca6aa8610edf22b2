// capi.cpp -- extern "C" boundary (include/macau_amd.h).
#include "../../include/macau_amd.h"

#include <cstring>
#include <string>
#include <vector>

#include "batch.h"
#include "model.h"

using namespace mgpu;

struct macau_model {
  std::unique_ptr<Model> m;
};

struct macau_batch {
  std::unique_ptr<ModelBatch> b;
};

static thread_local std::string g_err;

#define API_BEGIN try {
#define API_END_INT                      \
  }                                      \
  catch (const std::exception& e) {      \
    g_err = e.what();                    \
    return -1;                           \
  }                                      \
  catch (...) {                          \
    g_err = "unknown error";             \
    return -1;                           \
  }                                      \
  return 0;
#define API_END_VAL(v)                   \
  }                                      \
  catch (const std::exception& e) {      \
    g_err = e.what();                    \
    return v;                            \
  }                                      \
  catch (...) {                          \
    g_err = "unknown error";             \
    return v;                            \
  }

static Model& M(macau_model* h) {
  if (!h || !h->m) throw std::runtime_error("null model");
  return *h->m;
}

extern "C" {

const char* macau_last_error(void) { return g_err.c_str(); }
const char* macau_version(void) { return "0.1.0"; }

macau_model* macau_make_fixed(int nmodes, int K, double precision) {
  API_BEGIN
  auto* h = new macau_model;
  h->m.reset(new Model(nmodes, K, std::unique_ptr<INoiseModel>(new FixedGaussianNoise(precision))));
  return h;
  API_END_VAL(nullptr)
}

macau_model* macau_make_adaptive(int nmodes, int K, double sn_init, double sn_max) {
  API_BEGIN
  auto* h = new macau_model;
  h->m.reset(new Model(nmodes, K, std::unique_ptr<INoiseModel>(new AdaptiveGaussianNoise(sn_init, sn_max))));
  return h;
  API_END_VAL(nullptr)
}

macau_model* macau_make_probit(int nmodes, int K) {
  API_BEGIN
  auto* h = new macau_model;
  h->m.reset(new Model(nmodes, K, std::unique_ptr<INoiseModel>(new ProbitNoise())));
  return h;
  API_END_VAL(nullptr)
}

void macau_free(macau_model* h) { delete h; }

int macau_add_bpmf_prior(macau_model* h) {
  API_BEGIN
  M(h).addPrior(std::unique_ptr<ILatentPrior>(new BPMFPrior()));
  API_END_INT
}

int macau_add_macau_prior(macau_model* h, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                          const double* vals, double lambda_beta, double tol) {
  API_BEGIN
  std::vector<int> r(rows, rows + nnz), c(cols, cols + nnz);
  std::vector<double> v;
  if (vals) v.assign(vals, vals + nnz);
  M(h).addPrior(std::unique_ptr<ILatentPrior>(new MacauPrior(nrow, ncol, r, c, v, vals == nullptr, lambda_beta, tol)));
  API_END_INT
}

int macau_add_macau_prior_ff(macau_model* h, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                             const double* vals, double lambda_beta, double tol, int max_ff_size) {
  API_BEGIN
  std::vector<int> r(rows, rows + nnz), c(cols, cols + nnz);
  std::vector<double> v;
  if (vals) v.assign(vals, vals + nnz);
  auto* p = new MacauPrior(nrow, ncol, r, c, v, vals == nullptr, lambda_beta, tol);
  p->use_ftf = ncol <= max_ff_size;  // compute_ff (macau.pyx:160)
  M(h).addPrior(std::unique_ptr<ILatentPrior>(p));
  API_END_INT
}

int macau_add_macau_one_prior(macau_model* h, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                              const double* vals, double lambda_beta) {
  API_BEGIN
  std::vector<int> r(rows, rows + nnz), c(cols, cols + nnz);
  std::vector<double> v;
  if (vals) v.assign(vals, vals + nnz);
  M(h).addPrior(std::unique_ptr<ILatentPrior>(new MacauOnePrior(nrow, ncol, r, c, v, vals == nullptr, lambda_beta)));
  API_END_INT
}

int macau_set_samples(macau_model* h, int burnin, int nsamples) {
  API_BEGIN
  M(h).setSamples(burnin, nsamples);
  API_END_INT
}

int macau_set_relation_data(macau_model* h, const int* idx, int nmodes, const double* values, int nnz,
                            const int* dims) {
  API_BEGIN
  M(h).setRelationData(idx, nmodes, values, nnz, dims);
  API_END_INT
}

int macau_set_relation_data_test(macau_model* h, const int* idx, int nmodes, const double* values, int nnz,
                                 const int* dims) {
  API_BEGIN
  M(h).setRelationDataTest(idx, nmodes, values, nnz, dims);
  API_END_INT
}

int macau_set_verbose(macau_model* h, int v) {
  API_BEGIN
  M(h).verbose = v != 0;
  API_END_INT
}
int macau_set_save_model(macau_model* h, int save) {
  API_BEGIN
  M(h).save_model = save != 0;
  API_END_INT
}
int macau_set_save_prefix(macau_model* h, const char* prefix) {
  API_BEGIN
  M(h).save_prefix = prefix ? prefix : "model";
  API_END_INT
}
int macau_set_seed(macau_model* h, unsigned long long seed) {
  API_BEGIN
  M(h).seed = seed;
  API_END_INT
}
int macau_set_sampler(macau_model* h, int kind) {
  API_BEGIN
  if (kind != SAMPLER_PLAIN && kind != SAMPLER_BATCHED) throw std::runtime_error("unknown sampler kind");
  M(h).sampler = kind;
  API_END_INT
}
int macau_set_graph(macau_model* h, int on) {
  API_BEGIN
  M(h).capture_graph(on != 0);
  API_END_INT
}

int macau_run(macau_model* h) {
  API_BEGIN
  M(h).run();
  API_END_INT
}
int macau_init(macau_model* h) {
  API_BEGIN
  M(h).init();
  API_END_INT
}
int macau_iterate(macau_model* h, int n) {
  API_BEGIN
  M(h).iterate(n);
  API_END_INT
}
int macau_synchronize(macau_model* h) {
  API_BEGIN
  M(h).sync();
  M(h).check_fail();
  API_END_INT
}
int macau_profile(macau_model* h, int n, double* phase_ms) {
  API_BEGIN
  auto ms = M(h).profile(n);
  memcpy(phase_ms, ms.data(), ms.size() * sizeof(double));
  API_END_INT
}
int macau_sampler_supported(int kind, int num_latent, int nmodes) {
  if (num_latent < 1 || num_latent > 128 || nmodes < 2 || nmodes > kMaxOther + 1) return 0;
  if (kind == SAMPLER_PLAIN) return 1;
  if (kind == SAMPLER_BATCHED) return batched_supported(pad_latent_batched(num_latent), nmodes - 1);
  return 0;
}
int macau_get_iteration(macau_model* h) {
  API_BEGIN
  return M(h).host_iter;
  API_END_VAL(-1)
}

double macau_get_rmse_test(macau_model* h) {
  API_BEGIN
  return M(h).getRmseTest();
  API_END_VAL(NAN)
}
double macau_get_rmse_onesample(macau_model* h) {
  API_BEGIN
  return M(h).host_eval_out(0);
  API_END_VAL(NAN)
}
double macau_get_rmse_avg(macau_model* h) {
  API_BEGIN
  return M(h).host_eval_out(1);
  API_END_VAL(NAN)
}
int macau_get_auc(macau_model* h, double* auc_onesample, double* auc_avg) {
  API_BEGIN
  auto* p = dynamic_cast<ProbitNoise*>(M(h).noise.get());
  if (!p) throw std::runtime_error("AUC is reported by probit models only");
  if (!M(h).initialized) throw std::runtime_error("model not initialized");
  M(h).sync();
  M(h).check_fail();
  p->sync_host(M(h));
  if (auc_onesample) *auc_onesample = p->auc_test_onesample;
  if (auc_avg) *auc_avg = p->auc_test;
  API_END_INT
}

double macau_get_noise_precision(macau_model* h) {
  API_BEGIN
  return M(h).noise_precision();
  API_END_VAL(NAN)
}
double macau_get_mean_value(macau_model* h) {
  API_BEGIN
  return M(h).mean_value;
  API_END_VAL(NAN)
}
long long macau_get_test_nonzeros(macau_model* h) {
  API_BEGIN
  return M(h).ntest;
  API_END_VAL(-1)
}
int macau_get_dims(macau_model* h, int* dims) {
  API_BEGIN
  for (size_t i = 0; i < M(h).dims.size(); i++) dims[i] = M(h).dims[i];
  API_END_INT
}
int macau_get_predictions(macau_model* h, double* out) {
  API_BEGIN
  auto p = M(h).getPredictions();
  if (!p.empty()) memcpy(out, p.data(), p.size() * sizeof(double));
  API_END_INT
}
int macau_get_stds(macau_model* h, double* out) {
  API_BEGIN
  auto p = M(h).getStds();
  if (!p.empty()) memcpy(out, p.data(), p.size() * sizeof(double));
  API_END_INT
}
int macau_get_test_data(macau_model* h, double* out) {
  API_BEGIN
  auto p = M(h).getTestData();
  if (!p.empty()) memcpy(out, p.data(), p.size() * sizeof(double));
  API_END_INT
}
int macau_get_latents(macau_model* h, int mode, double* out) {
  API_BEGIN
  Model& m = M(h);
  if (mode < 0 || mode >= m.nmodes) throw std::runtime_error("mode out of range");
  std::vector<double> U;
  m.download_latents(mode, U);
  memcpy(out, U.data(), U.size() * sizeof(double));
  API_END_INT
}
int macau_get_hyper(macau_model* h, int mode, double* mu, double* Lambda) {
  API_BEGIN
  Model& m = M(h);
  if (mode < 0 || mode >= m.nmodes) throw std::runtime_error("mode out of range");
  const int K = m.num_latent;
  ILatentPrior& p = *m.priors[mode];
  HIPCHK(hipMemcpyAsync(mu, p.mu.as<void>(), K * sizeof(double), hipMemcpyDeviceToHost, m.stream));
  HIPCHK(hipMemcpyAsync(Lambda, p.Lambda.as<void>(), (size_t)K * K * sizeof(double), hipMemcpyDeviceToHost,
                        m.stream));
  m.sync();
  API_END_INT
}
int macau_get_link(macau_model* h, int mode, double* beta) {
  API_BEGIN
  Model& m = M(h);
  auto* p = dynamic_cast<MacauPrior*>(m.priors.at(mode).get());
  if (!p) throw std::runtime_error("mode has no side information");
  HIPCHK(hipMemcpyAsync(beta, p->beta.as<void>(), (size_t)m.num_latent * p->nfeat * sizeof(double),
                        hipMemcpyDeviceToHost, m.stream));
  m.sync();
  API_END_INT
}
double macau_get_link_lambda(macau_model* h, int mode) {
  API_BEGIN
  Model& m = M(h);
  return m.priors.at(mode)->getLinkLambda(m);
  API_END_VAL(NAN)
}
int macau_get_last_cg_iters(macau_model* h, int mode) {
  API_BEGIN
  Model& m = M(h);
  auto* p = dynamic_cast<MacauPrior*>(m.priors.at(mode).get());
  return p ? p->last_cg_iters : -1;
  API_END_VAL(-1)
}
int macau_get_ftf_solver(macau_model* h, int mode) {
  API_BEGIN
  Model& m = M(h);
  auto* p = dynamic_cast<MacauPrior*>(m.priors.at(mode).get());
  return p && p->use_ftf ? p->ftf_solver : 0;
  API_END_VAL(-1)
}

// ---------------------------------------------------------------- stateless
int macau_sample_latents(int K, int nother, int nrows, const long long* row_ptr, const int* idx, const double* vals,
                         double mean, const double* const* others, const int* other_n, double alpha,
                         const double* mu, const double* Lambda, const double* mu_rows, unsigned long long seed,
                         int iter, int mode, int sampler, double* U_out) {
  API_BEGIN
  if (K < 1 || K > 128) throw std::runtime_error("num_latent must be in [1, 128]");
  if (nother < 1 || nother > kMaxOther) throw std::runtime_error("nother must be in [1, 9]");
  const bool batched = sampler == SAMPLER_BATCHED || sampler == MACAU_SAMPLER_BATCHED_NOREFINE;
  const int KP = batched ? pad_latent_batched(K) : pad_latent(K);
  const long long nnz = row_ptr[nrows];
  if (nnz > 0x7fffffffLL) throw std::runtime_error("too many nonzeros for one mode");
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } guard{s};
  std::vector<int> rp(nrows + 1);
  for (int i = 0; i <= nrows; i++) rp[i] = (int)row_ptr[i];
  std::vector<int> ix(idx, idx + (size_t)nnz * nother);
  std::vector<float> r(nnz);
  for (long long e = 0; e < nnz; e++) {
    r[e] = (float)(vals[e] - mean);
    for (int t = 0; t < nother; t++)
      if (ix[(size_t)e * nother + t] < 0 || ix[(size_t)e * nother + t] >= other_n[t])
        throw std::runtime_error("index out of range of the other factor matrix");
  }
  DevBuf d_rp, d_ix, d_r, d_alpha, d_iter, d_fail, d_Lf, d_Lmu, d_Lmur, d_U;
  std::vector<DevBuf> d_oth(nother);
  upload(d_rp, rp, s);
  upload(d_ix, ix, s);
  upload(d_r, r, s);
  upload(d_alpha, std::vector<double>{alpha}, s);
  upload(d_iter, std::vector<int>{iter}, s);
  upload(d_fail, std::vector<int>{0}, s);
  FactorSet fs;
  for (int t = 0; t < nother; t++) {
    std::vector<float> f((size_t)KP * other_n[t], 0.f);
    for (int n = 0; n < other_n[t]; n++)
      for (int k = 0; k < K; k++) f[(size_t)n * KP + k] = (float)others[t][(size_t)n * K + k];
    upload(d_oth[t], f, s);
    HIPCHK(hipStreamSynchronize(s));
    fs.f[t] = d_oth[t].as<float>();
    fs.n[t] = other_n[t];
  }
  std::vector<float> Lf((size_t)KP * KP, 0.f), Lmu(KP, 0.f);
  for (int i = 0; i < KP; i++)
    for (int j = 0; j < KP; j++)
      Lf[(size_t)j * KP + i] = (i < K && j < K) ? (float)Lambda[(size_t)j * K + i] : (i == j ? 1.f : 0.f);
  for (int i = 0; i < K; i++) {
    double t = 0.0;
    for (int j = 0; j < K; j++) t += Lambda[(size_t)j * K + i] * mu[j];
    Lmu[i] = (float)t;
  }
  upload(d_Lf, Lf, s);
  upload(d_Lmu, Lmu, s);
  if (mu_rows) {
    std::vector<float> lm((size_t)KP * nrows, 0.f);
    for (int n = 0; n < nrows; n++)
      for (int i = 0; i < K; i++) {
        double t = 0.0;
        for (int j = 0; j < K; j++) t += Lambda[(size_t)j * K + i] * mu_rows[(size_t)n * K + j];
        lm[(size_t)n * KP + i] = (float)t;
      }
    upload(d_Lmur, lm, s);
    HIPCHK(hipStreamSynchronize(s));
  }
  d_U.alloc((size_t)KP * std::max(nrows, 1) * sizeof(float));
  SampleArgs a;
  a.rows.N = nrows;
  a.rows.nother = nother;
  a.rows.row_ptr = d_rp.as<int>();
  a.rows.idx = d_ix.as<int>();
  a.rows.r = d_r.as<float>();
  a.others = fs;
  a.alpha = d_alpha.as<double>();
  a.Lambda_f = d_Lf.as<float>();
  a.Lmu_f = d_Lmu.as<float>();
  a.Lmu_rows = mu_rows ? d_Lmur.as<float>() : nullptr;
  a.U = d_U.as<float>();
  a.K = K;
  a.KP = KP;
  a.seed = seed;
  a.iter = d_iter.as<int>();
  a.mode = mode;
  a.fail = d_fail.as<int>();
  std::vector<DevBuf> pb, qb;
  if (batched) {
    if (!batched_supported(KP, nother)) throw std::runtime_error("batched sampler: unsupported latent size");
    if (!batched_fits(fs, nother, KP)) throw std::runtime_error("batched sampler: a factor matrix exceeds 2^31 bytes");
    BatchPlan p = build_batch_plan(rp, KP, pb, s);
    if (sampler == SAMPLER_BATCHED && KP <= 64) a.refine = make_refine_queue(qb, KP, std::min(std::max(nrows, 1), kRefineCap),
                                                                 kRefineThresh, std::max(K, kRefineMinNnz), s);
    launch_sample_batched(a, p, s);
  } else if (sampler == SAMPLER_PLAIN) {
    launch_sample_plain(a, s);
  } else {
    throw std::runtime_error("unknown sampler kind");
  }
  HIPCHK(hipGetLastError());
  std::vector<float> U((size_t)KP * nrows);
  int fail = 0;
  HIPCHK(hipMemcpyAsync(U.data(), d_U.as<void>(), U.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&fail, d_fail.as<void>(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (fail) throw std::runtime_error("Cholesky Decomposition failed!");
  for (int n = 0; n < nrows; n++)
    for (int k = 0; k < K; k++) U_out[(size_t)n * K + k] = U[(size_t)n * KP + k];
  API_END_INT
}

int macau_cond_normal_wishart(int K, int N, const double* U, const double* mu0, double kappa, const double* T,
                              double nu, unsigned long long seed, int iter, int mode, double* mu_out,
                              double* Lambda_out) {
  API_BEGIN
  if (K < 1 || K > 128) throw std::runtime_error("num_latent must be in [1, 128]");
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } guard{s};
  const int KP = pad_latent(K);
  DevBuf dU, dmu0, dT, dpart, dws, dmu, dL, dLf, dLmu, diter, dfail;
  upload(dU, std::vector<double>(U, U + (size_t)K * N), s);
  upload(dmu0, std::vector<double>(mu0, mu0 + K), s);
  upload(dT, std::vector<double>(T, T + (size_t)K * K), s);
  upload(diter, std::vector<int>{iter}, s);
  upload(dfail, std::vector<int>{0}, s);
  const int nblk = hyper_stats_blocks(N);
  dpart.alloc((size_t)nblk * (K + K * K) * sizeof(double));
  if (nw_workspace_bytes(K)) dws.alloc(nw_workspace_bytes(K));
  dmu.alloc(K * sizeof(double));
  dL.alloc((size_t)K * K * sizeof(double));
  dLf.alloc((size_t)KP * KP * sizeof(float));
  dLmu.alloc(KP * sizeof(float));
  launch_hyper_stats_f64(dU.as<double>(), K, N, K, dpart.as<double>(), nblk, s);
  NWArgs a;
  a.K = K;
  a.KP = KP;
  a.partial = dpart.as<double>();
  a.nblk = 1;  // totals in row 0
  a.N = N;
  a.kappa = kappa;
  a.mu0 = dmu0.as<double>();
  a.WI = dT.as<double>();
  a.nu = nu;
  a.seed = seed;
  a.iter = diter.as<int>();
  a.mode = mode;
  a.ws = dws.as<double>();
  a.mu = dmu.as<double>();
  a.Lambda = dL.as<double>();
  a.Lambda_f = dLf.as<float>();
  a.Lmu_f = dLmu.as<float>();
  a.fail = dfail.as<int>();
  launch_nw_draw(a, s);
  HIPCHK(hipGetLastError());
  int fail = 0;
  HIPCHK(hipMemcpyAsync(mu_out, dmu.as<void>(), K * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(Lambda_out, dL.as<void>(), (size_t)K * K * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&fail, dfail.as<void>(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (fail) throw std::runtime_error("CondNormalWishart: matrix not positive definite");
  API_END_INT
}

int macau_set_refine(macau_model* m, int on, double threshold) {
  API_BEGIN
  Model& x = M(m);
  x.refine = on != 0;
  if (threshold > 0) x.refine_thresh = (float)threshold;
  API_END_INT
}

int macau_get_refine_stats(macau_model* m, long long* flagged, long long* refined) {
  API_BEGIN
  Model& x = M(m);
  long long t[2] = {0, 0};
  int c = 0;
  if (x.rq.count) {
    HIPCHK(hipStreamSynchronize(x.stream));
    HIPCHK(hipMemcpy(t, x.rq.totals, sizeof(t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&c, x.rq.count, sizeof(int), hipMemcpyDeviceToHost));
  }
  // the refinement launch folds each half-sweep's queue into the totals (c is
  // nonzero only while a half-sweep is in flight)
  if (flagged) *flagged = t[0] + c;
  if (refined) *refined = t[1] + std::min(c, x.rq.cap);
  API_END_INT
}

int macau_autotune(macau_model* m, int reps, const char* record_path, double* default_ms, double* tuned_ms) {
  API_BEGIN
  auto r = M(m).autotune(reps, record_path ? std::string(record_path) : std::string());
  for (size_t q = 0; q < r.size(); q++) {
    if (default_ms) default_ms[q] = r[q].first;
    if (tuned_ms) tuned_ms[q] = r[q].second;
  }
  API_END_INT
}

int macau_get_tuning(macau_model* m, int mode, int* knobs) {
  API_BEGIN
  Model& x = M(m);
  if (!x.initialized) throw std::runtime_error("model not initialized");
  if (mode < 0 || mode >= x.nmodes) throw std::runtime_error("mode out of range");
  const TuneChoice& c = x.tune[mode];
  knobs[0] = c.seg_len;
  knobs[1] = c.light_cap ? c.light_cap : kDefaultLightCap;
  knobs[2] = c.heavy_cap ? c.heavy_cap : kDefaultHeavyCap;
  knobs[3] = c.seg_cap ? c.seg_cap : kDefaultSegCap;
  return c.tuned ? 1 : 0;
  API_END_VAL(-1)
}

macau_batch* macau_batch_create(macau_model* const* models, int n) {
  API_BEGIN
  if (n < 1 || !models) throw std::runtime_error("a batch needs at least one model");
  std::vector<Model*> ms;
  for (int q = 0; q < n; q++) ms.push_back(&M(models[q]));
  auto* h = new macau_batch;
  h->b.reset(new ModelBatch(ms));
  return h;
  API_END_VAL(nullptr)
}

void macau_batch_free(macau_batch* b) { delete b; }

static ModelBatch& B(macau_batch* h) {
  if (!h || !h->b) throw std::runtime_error("null batch");
  return *h->b;
}

int macau_batch_set_graph(macau_batch* b, int on) {
  API_BEGIN
  B(b).use_graph = on != 0;
  API_END_INT
}

int macau_batch_iterate(macau_batch* b, int n) {
  API_BEGIN
  B(b).iterate(n);
  API_END_INT
}

int macau_batch_synchronize(macau_batch* b) {
  API_BEGIN
  B(b).synchronize();
  API_END_INT
}

int macau_cond_normal_wishart_noalloc(int num_latent, int N, const double* d_U, const double* d_mu0, double kappa,
                                      const double* d_T, double nu, unsigned long long seed, int iter, int mode,
                                      double* d_mu_out, double* d_Lambda_out, int* d_status, void* stream) {
  API_BEGIN
  cond_normal_wishart_noalloc(num_latent, N, d_U, d_mu0, kappa, d_T, nu, seed, iter, mode, d_mu_out, d_Lambda_out,
                              d_status, static_cast<hipStream_t>(stream));
  API_END_INT
}

int macau_solve_blockcg(int nrhs, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                        const double* vals, double reg, const double* B, double tol, double* X, int* iters) {
  API_BEGIN
  if (nrhs < 1 || nrhs > 40) throw std::runtime_error("BlockCG only available for up to 40 RHSs.");
  MacauPrior F(nrow, ncol, std::vector<int>(rows, rows + nnz), std::vector<int>(cols, cols + nnz),
               vals ? std::vector<double>(vals, vals + nnz) : std::vector<double>(), vals == nullptr, reg, tol);
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } guard{s};
  F.upload_side(nrhs, s);
  DevBuf dB, dX, dreg;
  upload(dB, std::vector<double>(B, B + (size_t)nrhs * ncol), s);
  upload(dreg, std::vector<double>{reg}, s);
  dX.alloc((size_t)nrhs * ncol * sizeof(double));
  CgBuffers cg;
  cg.alloc(nrhs, ncol, nrow);
  HIPCHK(hipStreamSynchronize(s));
  const CsrDev fd = F.Fd(), ftd = F.Ftd();
  const int it = blockcg_device(F.F(), F.Ft(), dreg.as<double>(), dB.as<double>(), nrhs, 0, nrhs, tol,
                                dX.as<double>(), cg, s, false, &fd, &ftd);
  HIPCHK(hipMemcpyAsync(X, dX.as<void>(), (size_t)nrhs * ncol * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (iters) *iters = it;
  API_END_INT
}

int macau_rng_probe(unsigned long long seed, unsigned c0, unsigned iter, unsigned tag, int n, unsigned* bits,
                    float* z32, double* z64) {
  API_BEGIN
  DevBuf db((size_t)n * 4), d32((size_t)n * 4), d64((size_t)n * 8);
  launch_rng_probe(seed, c0, iter, tag, n, db.as<uint32_t>(), d32.as<float>(), d64.as<double>(), nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(bits, db.as<void>(), (size_t)n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(z32, d32.as<void>(), (size_t)n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(z64, d64.as<void>(), (size_t)n * 8, hipMemcpyDeviceToHost));
  API_END_INT
}

int macau_truncnorm_probe(unsigned long long seed, unsigned c0, unsigned iter, unsigned tag, int n,
                          const double* mean, const double* p, double* out, double* inv) {
  API_BEGIN
  DevBuf dm((size_t)n * 8), dp((size_t)n * 8), d((size_t)n * 8), di((size_t)n * 8);
  HIPCHK(hipMemcpy(dm.as<void>(), mean, (size_t)n * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dp.as<void>(), p, (size_t)n * 8, hipMemcpyHostToDevice));
  launch_truncnorm_probe(n, dm.as<double>(), dp.as<double>(), seed, c0, iter, tag, d.as<double>(),
                         di.as<double>(), nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, d.as<void>(), (size_t)n * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(inv, di.as<void>(), (size_t)n * 8, hipMemcpyDeviceToHost));
  API_END_INT
}

int macau_gamma_probe(unsigned long long seed, const double* shapes, int n, unsigned iter, unsigned tag,
                      double* out) {
  API_BEGIN
  DevBuf ds((size_t)n * 8), d((size_t)n * 8);
  HIPCHK(hipMemcpy(ds.as<void>(), shapes, (size_t)n * 8, hipMemcpyHostToDevice));
  launch_gamma_probe(seed, ds.as<double>(), n, iter, tag, d.as<double>(), nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, d.as<void>(), (size_t)n * 8, hipMemcpyDeviceToHost));
  API_END_INT
}

}  // extern "C"
