// model.h -- host C++ mirror of the reference's Macau / ILatentPrior /
// INoiseModel interfaces (macau.h:11-81, latentprior.h:17-125,
// noisemodels.h:13-86), backed by device-resident state on one MI355X.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace mgpu {

void hip_check(hipError_t e, const char* what);
#define HIPCHK(x) ::mgpu::hip_check((x), #x)

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(size_t bytes);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p_); }
  size_t bytes() const { return n_; }
 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

enum SamplerKind { SAMPLER_PLAIN = 0, SAMPLER_BATCHED = 1 };

// launch shape of the batched sampler for one mode (tune.cpp); 0 caps = defaults
struct TuneChoice {
  int seg_len = kDefaultSegLen, light_cap = 0, heavy_cap = 0, seg_cap = 0;
  bool tuned = false;               // chosen by the autotuner (now or from the table)
  double default_ms = 0, ms = 0;    // measured half-sweep times (default shape, chosen shape)
};

// Sparse rows of one mode on the host (SparseMode, sparsetensor.h:76-91).
struct HostRows {
  std::vector<int> row_ptr;  // N + 1
  std::vector<int> idx;      // nnz * nother
  std::vector<double> vals;  // raw values
};

class Model;

// Buffers of one block-CG solve (solve_blockcg, linop.h:228-343).
struct CgBuffers {
  DevBuf X, R, P, KP, B, inner, partial, small;
  int nrmax = 0;
  int last_iters = 0;  // iterations of the previous solve (sizes the first chunk)
  void alloc(int nrmax, int nfeat, int nsamples);
  // instantiated graphs of a chunk of CG iterations, one per launch
  // configuration (the kernels' arguments: operators, buffers, regulariser,
  // tolerance) and chunk length
  struct Chunk {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Chunk> chunks;
  CgBuffers() = default;
  CgBuffers(const CgBuffers&) = delete;
  CgBuffers& operator=(const CgBuffers&) = delete;
  ~CgBuffers();
};
// Solves X[r0:r0+nr, :] (F^T F + reg I) = B[r0:r0+nr, :] with B and X of
// leading dimension ld (= total right-hand sides); reg is a device scalar.
// Returns the iteration count of the reference's loop.  With use_graph the
// chunks of CG iterations launched per host check of the convergence flags
// are captured once per length and replayed as HIP graphs.
int blockcg_device(const CsrDev& F, const CsrDev& Ft, const double* reg, const double* B, int ld, int r0, int nr,
                   double tol, double* X, CgBuffers& b, hipStream_t s, bool use_graph = false,
                   const CsrDev* Fd = nullptr, const CsrDev* Ftd = nullptr);

// Work list of the batched sampler for one mode (rows sorted by nnz,
// heavy rows split into segments).
BatchPlan build_batch_plan(const std::vector<int>& row_ptr, int KP, std::vector<DevBuf>& bufs, hipStream_t s,
                           int seg_len = kDefaultSegLen);

// latentprior.h:17-49
class ILatentPrior {
 public:
  virtual ~ILatentPrior() = default;
  virtual void init(Model& m, int mode);
  virtual void sample_latents(Model& m, int mode);
  virtual void update_prior(Model& m, int mode) = 0;
  virtual double getLinkNorm(Model&) { return NAN; }
  virtual double getLinkLambda(Model&) { return NAN; }
  virtual void saveModel(Model& m, int mode, const std::string& prefix);
  virtual const float* prior_mean_rows() const { return nullptr; }

  // device parameters shared with the sampler (FP64 state + FP32 copies)
  DevBuf mu, Lambda, Lambda_f, Lmu_f, mu0, WI, stats, nw_ws;
  int stats_nblk = 0;
};

// BPMFPrior (latentprior.h:53-79, latentprior.cpp:69-99)
class BPMFPrior : public ILatentPrior {
 public:
  void update_prior(Model& m, int mode) override;
};

constexpr int kFtfCholesky = 1, kFtfSpectral = 2;

// MacauPrior<SparseDoubleFeat> (latentprior.h:82-125, latentprior.cpp:180-331)
class MacauPrior : public ILatentPrior {
 public:
  MacauPrior(int nrow, int ncol, const std::vector<int>& rows, const std::vector<int>& cols,
             const std::vector<double>& vals, bool binary, double lambda_beta, double tol);
  void init(Model& m, int mode) override;
  void update_prior(Model& m, int mode) override;
  double getLinkNorm(Model& m) override;
  double getLinkLambda(Model& m) override;
  void saveModel(Model& m, int mode, const std::string& prefix) override;
  const float* prior_mean_rows() const override { return Lmu_rows.as<float>(); }
  // whether update_prior can be captured in a HIP graph: the block CG reads
  // its convergence flags on the host after each chunk of iterations; the F^T F branch
  // launches a fixed sequence of kernels
  virtual bool capturable() const { return use_ftf; }

  int nrow, nfeat;
  double lambda_beta0, tol;
  bool binary;
  std::vector<int> F_rowptr, F_cols, Ft_rowptr, Ft_cols;
  std::vector<double> F_vals, Ft_vals;
  DevBuf dF_rowptr, dF_cols, dF_vals, dFt_rowptr, dFt_cols, dFt_vals;
  DevBuf beta, Uhat, Lmu_rows, lambda_beta_d, Lchol, tmp, Fty, bbt, partial_sumsq;
  CgBuffers cg;
  int bbt_nblk = 0;
  int last_cg_iters = 0;
  // F^T F branch of sample_beta (latentprior.cpp:280-289): set for side
  // information with at most max_ff_size features (macau.pyx:160)
  bool use_ftf = false;
  // F^T F (lower, FP64) and the per-iteration Cholesky work matrix; with the
  // spectral solve they hold the eigenvectors Q and Q^T instead
  DevBuf FtF, Wchol;
  // F^T F solver: kFtfCholesky, or kFtfSpectral after a successful
  // eigendecomposition at init (MACAU_AMD_FTF_SOLVER=cholesky forces the former)
  int ftf_solver = 0;
  std::string ftf_note;  // why the spectral solve is not used, when it is not
  DevBuf ftf_d, ftf_y, ftf_part;
  SpectralPlan ftf_plan;
  CsrDev F() const;
  CsrDev Ft() const;
  // uploads F and F^T and builds their SpMM work lists (nrmax right-hand sides)
  void upload_side(int nrmax, hipStream_t s);
  CsrDev F_plan, Ft_plan;
  std::vector<DevBuf> F_plan_bufs, Ft_plan_bufs;
  // the same products with a dense column block (binary F with at least 16
  // columns holding a nonzero in every 64th row or more; 32 right-hand
  // sides): dense_role 0 when there is none
  CsrDev Fd() const;
  CsrDev Ftd() const;
  void build_dense(int nrmax, hipStream_t s);
  CsrDev Fd_plan, Ftd_plan;
  std::vector<DevBuf> dense_bufs, Ftd_plan_bufs;
};

// MacauOnePrior<SparseDoubleFeat> (macauoneprior.h:8-51): diagonal prior
// precision, coordinate-wise row updates (the plain sampler's univariate
// variant) and a feature-sequential link-matrix sweep per latent.  Shares
// the side-information storage, beta, Uhat and the per-row prior term of
// MacauPrior.
class MacauOnePrior : public MacauPrior {
 public:
  bool capturable() const override { return true; }  // a fixed sequence of kernels
  MacauOnePrior(int nrow, int ncol, const std::vector<int>& rows, const std::vector<int>& cols,
                const std::vector<double>& vals, bool binary, double lambda_beta);
  void init(Model& m, int mode) override;
  void sample_latents(Model& m, int mode) override;
  void update_prior(Model& m, int mode) override;
  double getLinkLambda(Model& m) override;  // lambda_beta.mean() (macauoneprior.h:45)

  std::vector<double> colsq;           // col_square_sum(F) (linop.cpp:374-401)
  std::vector<int> Fu_rowptr, Fu_cols;  // F^T with repeated entries of a row merged
  std::vector<double> Fu_vals;
  std::vector<int4> batch_meta;  // the sweep's batches of features with disjoint columns
  DevBuf dFu_rowptr, dFu_cols, dFu_vals, dcolsq, lambda_beta_vec, uhat64, ainv, nz, Zw, dbatch, lb_partial;
};

// noisemodels.h:13-29
class INoiseModel {
 public:
  virtual ~INoiseModel() = default;
  virtual void init(Model& m) = 0;
  virtual void update(Model&) {}
  virtual void evalModel(Model& m);
  virtual double getEvalMetric() { return rmse_test; }
  virtual std::string getEvalString();
  virtual std::string getInitStatus() = 0;
  virtual std::string getStatus() = 0;
  virtual void sync_host(Model&) {}
  // enqueued before mode `mode` is sampled (probit data augmentation)
  virtual void augment(Model&, int) {}
  double rmse_test = NAN, rmse_test_onesample = NAN;
};

// FixedGaussianNoise (noisemodels.h:32-50)
class FixedGaussianNoise : public INoiseModel {
 public:
  explicit FixedGaussianNoise(double a) : alpha(a) {}
  void init(Model& m) override;
  std::string getInitStatus() override;
  std::string getStatus() override { return ""; }
  double alpha;
};

// AdaptiveGaussianNoise (noisemodels.h:53-86, noisemodels.cpp:14-124)
class AdaptiveGaussianNoise : public INoiseModel {
 public:
  AdaptiveGaussianNoise(double sinit, double smax) : sn_init(sinit), sn_max(smax) {}
  void init(Model& m) override;
  void update(Model& m) override;
  std::string getInitStatus() override;
  std::string getStatus() override;
  void sync_host(Model& m) override;
  double sn_init, sn_max;
  double alpha = NAN, alpha_max = NAN, var_total = NAN;
};

// ProbitNoise (noisemodels.h:88-103, noisemodels.cpp:129-159, matrix data):
// precision 1, observations replaced by z before each mode is sampled;
// the test AUCs are computed on the host from the device predictions when
// sync_host is called.
class ProbitNoise : public INoiseModel {
 public:
  void init(Model& m) override;
  void augment(Model& m, int mode) override;
  void evalModel(Model& m) override;
  double getEvalMetric() override { return auc_test; }
  std::string getEvalString() override;
  std::string getInitStatus() override { return "Probit noise model"; }
  std::string getStatus() override { return ""; }
  void sync_host(Model& m) override;
  double auc_test = NAN, auc_test_onesample = NAN;
  std::vector<DevBuf> d_y, d_row_of, d_z;  // per mode
  DevBuf d_pred1;
};

// auc (bpmfutils.h:182-213): pred sorted ascending, ties by index
double auc(const std::vector<double>& pred, const std::vector<double>& test);

// MacauX<DType, NType> (macau.h:56-81, macau.cpp)
class Model {
 public:
  Model(int nmodes, int num_latent, std::unique_ptr<INoiseModel> noise);
  ~Model();

  // macau.h:35-53
  void addPrior(std::unique_ptr<ILatentPrior> p) { priors.push_back(std::move(p)); }
  void setSamples(int b, int n) { burnin = b; nsamples = n; }
  void setRelationData(const int* idx, int nmodes, const double* vals, int nnz, const int* dims);
  void setRelationDataTest(const int* idx, int nmodes, const double* vals, int nnz, const int* dims);
  void init();
  void run();
  void printStatus(int i, double elapsedi);
  double getRmseTest() const { return rmse_test; }
  std::vector<double> getPredictions();
  std::vector<double> getStds();
  std::vector<double> getTestData();  // ntest x (nmodes + 1), column-major
  void saveModel(int isample);
  void saveGlobalParams();

  // extensions
  void iterate(int n);                 // n Gibbs iterations after init()
  void sync() { HIPCHK(hipStreamSynchronize(stream)); }
  void check_fail();
  void download_latents(int mode, std::vector<double>& out);  // K x N col-major
  double noise_precision();
  void set_iteration(int it);
  void capture_graph(bool on) { use_graph = on; }
  // n iterations launched without a graph, with HIP events on the model's
  // stream around each phase; returns mean ms per phase:
  // [sample mode 0..M-1, prior updates (all modes), noise, eval]
  std::vector<double> profile(int n);

  // -- configuration --
  int nmodes, num_latent, KP;
  int burnin = 50, nsamples = 100;
  bool verbose = true, save_model = false;
  std::string save_prefix = "model";
  uint64_t seed = 0;
  int sampler = SAMPLER_PLAIN;
  bool use_graph = true;
  bool is_tensor;
  std::vector<std::unique_ptr<ILatentPrior>> priors;
  std::unique_ptr<INoiseModel> noise;
  double rmse_test = 0.0;

  // -- host data (sparsetensor.cpp) --
  std::vector<int> dims;
  std::vector<HostRows> rows;  // per mode
  double mean_value = 0.0;
  long long nnz = 0;
  bool has_test = false;
  std::vector<int> test_idx;  // ntest x nmodes
  std::vector<double> test_vals;
  long long ntest = 0;
  // device order of the test cells: device cell t is test cell test_perm[t]
  std::vector<long long> test_perm;
  // a per-test-cell device vector (d_pred, d_pred_var, ...) in the caller's order
  std::vector<double> test_to_host(const DevBuf& d);

  // -- device state --
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // prior updates overlapped with sampling (step_launch)
  std::vector<hipEvent_t> ev_fork;
  hipEvent_t ev_join = nullptr;
  hipEvent_t ev_start = nullptr, ev_plast = nullptr;  // deferred last-mode prior update (step_launch)
  hipEvent_t ev_noise = nullptr;  // the noise draw forked beside the evaluation
  // the batched sampler's heavy-row segment kernels, beside its light rows
  hipStream_t seg_stream = nullptr;
  std::vector<hipEvent_t> ev_seg_fork, ev_seg_join;
  std::vector<DevBuf> d_rowptr, d_idx, d_r, d_U;
  std::vector<RowsDev> rdev;
  std::vector<BatchPlan> plans;
  std::vector<std::vector<DevBuf>> plan_bufs;
  // d_iter keys the sampling, noise and evaluation draws; d_iter_prior the
  // prior updates' (it advances after the last mode's prior update, which
  // can run after d_iter has advanced: see step_launch)
  DevBuf d_iter_prior;
  DevBuf d_iter, d_alpha, d_fail, d_eval_out, d_test_idx, d_test_vals, d_pred, d_pred_var, d_eval_partial,
      d_noise_partial, d_norm_partial, d_sse, d_r2_last, d_sse_partial;
  int sse_partial_n = 0;       // fused SSE partials of the batched sampler (last mode)
  bool fused_sse = false;
  int eval_nblk = 0, noise_nblk = 0;
  bool initialized = false;
  int host_iter = 0;  // iterations launched so far
  // iteration graphs: [0] without, [1] with the previous iteration's
  // deferred last-mode prior update at its start
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  bool prior_pending = false;  // the last mode's prior update of the last launched iteration is deferred
  void drop_graphs();

  FactorSet others_of(int mode) const;
  SampleArgs sample_args(int mode, const ILatentPrior& p);
  // batched sampler: one refinement step for rows flagged by the probe
  // condition estimate (batched.h, refine.hip); settings read at init
  bool refine = true;
  float refine_thresh = kRefineThresh;
  int refine_min_nnz = -1;  // < 0: max(num_latent, kRefineMinNnz)
  std::vector<DevBuf> refine_bufs;
  RefineQueue rq;
  // batched sampler launch shapes (tune.cpp): per mode, looked up at init
  std::vector<TuneChoice> tune;
  std::string tune_key(int mode) const;
  void apply_cached_tuning();
  void apply_plan_shape(int mode);
  void update_sse_slots();
  // times the candidate shapes of every mode, keeps the fastest, returns
  // (default ms, chosen ms) per mode and appends the choices to record_path
  std::vector<std::pair<double, double>> autotune(int reps, const std::string& record_path);
  // enqueue one Gibbs iteration; defer_last leaves the last mode's prior
  // update for the next launch (prior_pending), run_pending runs the one
  // deferred by the previous launch at the start of this one
  void step_launch(bool defer_last = false, bool run_pending = false);
  void launch_pending_prior();  // enqueue the deferred prior update, if any
  double factor_norm(int mode);
  double host_eval_out(int which);

 private:
  void build_plans();
};

// Device queue of rows to refine, sized for cap rows of KP x KP factors
// (bufs receives the allocations).
RefineQueue make_refine_queue(std::vector<DevBuf>& bufs, int KP, int cap, float thresh, int min_nnz, hipStream_t s);

// Setup-time copy of a host vector into a new allocation; synchronous, so
// the host vector may be a temporary.
template <typename T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.as<void>(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
}

// CondNormalWishart on caller-owned device memory with no device allocation
// (noalloc.hip): U, mu0, T, mu_out, Lambda_out and status are device
// pointers; K <= 64; asynchronous on s
void cond_normal_wishart_noalloc(int K, int N, const double* U, const double* mu0, double kappa, const double* T,
                                 double nu, uint64_t seed, int iter, int mode, double* mu_out, double* Lambda_out,
                                 int* status, hipStream_t s);

}  // namespace mgpu
