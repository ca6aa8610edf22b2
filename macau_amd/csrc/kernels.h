// kernels.h -- host-side launch wrappers for the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace mgpu {

constexpr int kMaxOther = 9;  // up to 10-mode tensors (the reference's Data takes 10 index columns)
// kernels are compiled for 1..4 other modes; 5..kMaxOther use one variant
// compiled for kMaxOther with the count read at run time
constexpr int kMaxBatchedOther = 4;  // the batched sampler: up to 5-mode tensors

// Observed cells of one mode, grouped by that mode's index (the reference's
// SparseMode / Eigen CSC column view, sparsetensor.h:76-91).
struct RowsDev {
  int N = 0;                  // rows of this mode
  int nother = 1;             // other modes
  const int* row_ptr = nullptr;  // N + 1
  const int* idx = nullptr;      // nnz * nother, row-major, view order
  const float* r = nullptr;      // nnz: y - mean, rounded to FP32
};

struct FactorSet {
  const float* f[kMaxOther] = {};  // KP x N_m
  int n[kMaxOther] = {};           // N_m
};

// Work list of the batched sampler (rows sorted by nnz, heavy rows split).
constexpr int kDefaultSegLen = 2048;
constexpr int kDefaultLightCap = 8192, kDefaultHeavyCap = 2048, kDefaultSegCap = 16384;
// largest light / heavy caps the autotuner tries (SSE partial slots are sized for them)
constexpr int kMaxLightCap = 16384, kMaxHeavyCap = 4096;

struct BatchPlan {
  int KP = 0;                  // padded latent size the plan's partials are sized for
  int n_light = 0;             // light rows (sorted by nnz, descending)
  const int* light = nullptr;  // row ids
  int n_seg = 0;               // segments of heavy rows
  const int* seg_row = nullptr;    // heavy-row slot of each segment
  const int* seg_beg = nullptr;    // first nnz
  const int* seg_end = nullptr;    // one past last nnz
  int n_heavy = 0;
  const int* heavy = nullptr;      // row id of each heavy slot
  const int* heavy_seg0 = nullptr; // first segment of each heavy slot (n_heavy + 1)
  float* partial = nullptr;        // per-segment Gram (KP*KP) and rhs (KP)
  // light rows as the finishing kernel takes them: pairs of consecutive
  // light rows (KP = 32) or single rows (KP = 64, row y = -1)
  int n_pair = 0;
  const int4* pair_rng = nullptr;  // nnz ranges [x, y) and [z, w) (z = w = 0: no second row)
  const int2* pair_row = nullptr;  // row ids (y = -1: no second row)
  // launch shape (0 = the defaults below); chosen by the autotuner
  int seg_len = kDefaultSegLen;    // nonzeros per heavy-row segment (set when the plan is built)
  int light_cap = 0;               // most work-groups of the light-row kernel (default 8192)
  int heavy_cap = 0;               // most work-groups of the heavy-row finishing kernel (default 2048)
  int seg_cap = 0;                 // most work-groups of the segment kernel (default 16384)
};

// Rows of a batched half-sweep queued for one refinement step (refine.hip):
// a row whose probe condition estimate (batched.h, probe_estimate) exceeds
// thresh and which has at least min_nnz nonzeros is written here with its
// FP32 factor.
struct RefineQueue {
  int* count = nullptr;         // [0] rows queued in the current half-sweep (null: no refinement),
                                // [1] work-groups of the refinement launch that have finished
  int cap = 0;                  // entries
  int* row = nullptr;           // cap
  float* L = nullptr;           // cap x KP x KP, row-major, zero above the diagonal
  float thresh = 0.f;
  int min_nnz = 0;
  long long* totals = nullptr;  // [0] rows flagged, [1] rows refined, over the chain
  // short queues: per-chunk data terms (kRefineGrid x KP, FP64) and per-row
  // tickets (kRefineGrid, zero between launches) of the split rows
  double* part = nullptr;
  int* row_ticket = nullptr;
};
constexpr int kRefineGrid = 256;
constexpr int kRefineCap = 4096;
// default threshold of the probe condition estimate (an estimate of the
// condition number of the Jacobi-scaled P; K for a diagonal P): measured
// split-BF16 row errors stay below 5e-5 up to kappa ~ 500
constexpr float kRefineThresh = 500.f;
// rows with fewer nonzeros are neither probed nor queued (their errors
// stay small: the data term's rounding is spread over few products)
constexpr int kRefineMinNnz = 128;

struct SampleArgs {
  RowsDev rows;
  FactorSet others;
  const double* alpha = nullptr;   // device scalar: noise precision
  const float* Lambda_f = nullptr; // KP x KP, pad diagonal = 1
  const float* Lmu_f = nullptr;    // KP: Lambda mu (shared by all rows)
  const float* Lmu_rows = nullptr; // KP x N: Lambda (mu + Uhat_n), or null
  float* U = nullptr;              // KP x N, written
  int K = 0, KP = 0;
  uint64_t seed = 0;
  const int* iter = nullptr;       // device iteration counter
  int mode = 0;
  int* fail = nullptr;             // device flag: row + 1 of a failed Cholesky
  // algebraic train SSE of this half-sweep (batched sampler only), or null
  double* sse_partial = nullptr;
  const float* r2 = nullptr;       // per-row sum of squared residuals
  // univariate prior (MacauOnePrior, plain sampler only): Lambda_f is
  // diagonal and each row takes one coordinate-wise (Gauss-Seidel) sweep
  // from its current value instead of an exact draw
  int univariate = 0;
  // batched sampler: the refinement queue (count == null: no refinement)
  RefineQueue refine;
};
// refinement of the queued rows (refine.hip), launched after a batched
// half-sweep's epilogues; it leaves the queue cleared for the next half-sweep
void launch_refine(const SampleArgs& a, hipStream_t s);

void launch_sample_plain(const SampleArgs& a, hipStream_t s);
// With seg_stream set, the heavy-row kernels (segments, their sums and the
// heavy rows' finish) run on it beside the light rows: seg_fork is recorded
// on s before them and seg_join after them, and s waits on seg_join before
// the refinement step.
struct SegStreams {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
void launch_sample_batched(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg = {});
// K in (32, 64] (KP = 64): one row per wave, three 32 x 32 Gram blocks
void launch_sample_batched64(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg = {});
// K in (64, 128] (KP = 128): one row per 256-thread work-group, ten 32 x 32
// Gram blocks over four waves; no refinement queue (a.refine is ignored)
void launch_sample_batched128(const SampleArgs& a, const BatchPlan& p, hipStream_t s, const SegStreams& seg = {});
int batched64_light_grid(const BatchPlan& p);
int batched64_heavy_grid(const BatchPlan& p);
int batched_supported(int KP, int nother);
// the batched sampler addresses each other mode's factors with 32-bit byte
// offsets: N_m * KP * 4 must stay below 2^31
bool batched_fits(const FactorSet& f, int nother, int KP);
// grids of the batched finishing kernels; their SSE partials are laid out
// [light grid | heavy grid] when SampleArgs::sse_partial is set
int batched_light_grid(const BatchPlan& p);
int batched_heavy_grid(const BatchPlan& p);
int batched_seg_grid(const BatchPlan& p);

// --- hyperprior (FP64) ---
// sums of x = U[:,n] - sub[:,n] (sub may be null): sum (K) and scatter
// (K x K, col-major), accumulated per block into partial[nblk][K + K*K] and
// reduced in place: the totals are in partial[0] when the launch completes.
// hyper_stats_blocks(N) is the block count to allocate and pass.
int hyper_stats_blocks(int N);
void launch_hyper_stats_f32(const float* U, const float* sub, int ld, int N, int K, double* partial, int nblk,
                            hipStream_t s);
void launch_hyper_stats_f64(const double* U, int ld, int N, int K, double* partial, int nblk, hipStream_t s);

struct NWArgs {
  int K = 0, KP = 0;
  const double* partial = nullptr;  // nblk x (K + K*K), summed over the rows
  int nblk = 0;
  double N = 0;                     // number of columns
  double kappa = 2.0;
  const double* mu0 = nullptr;      // K
  const double* WI = nullptr;       // K x K base scale
  const double* BBt_partial = nullptr;  // Macau: partial stats of beta (or null)
  int bbt_nblk = 0;
  const double* lambda_beta = nullptr;  // device scalar (Macau)
  double nu = 0;                    // degrees of freedom
  uint64_t seed = 0;
  const int* iter = nullptr;
  int mode = 0;
  double* ws = nullptr;             // workspace (nw_workspace_bytes), or null for LDS
  double* mu = nullptr;             // out K
  double* Lambda = nullptr;         // out K x K
  double* Lchol = nullptr;          // out chol(Lambda), K x K (or null)
  float* Lambda_f = nullptr;        // out KP x KP
  float* Lmu_f = nullptr;           // out KP
  int* fail = nullptr;
};
void launch_nw_draw(const NWArgs& a, hipStream_t s);
// draws of several modes (same K) in one launch, one work-group per mode
struct NWBatch {
  NWArgs m[kMaxOther + 1];
  int n = 0;
};
void launch_nw_draw_batch(const NWBatch& b, hipStream_t s);
// global workspace the draw needs for latent size K (0: fits in LDS)
size_t nw_workspace_bytes(int K);

// --- noise and evaluation ---
struct NoiseArgs {
  RowsDev rows0;      // mode-0 rows
  const float* U0 = nullptr;
  FactorSet others;
  int K = 0, KP = 0, nmodes = 2;
  double* partial = nullptr;  // per-block SSE
  int nblk = 0;
};
int noise_sse_blocks(int N0);
void launch_noise_sse(const NoiseArgs& a, hipStream_t s);
struct NoiseDrawArgs {
  const double* partial = nullptr;
  int nblk = 0;
  double nnz = 0, var_total = 1, b0 = 0.5, alpha_max = 0;
  uint64_t seed = 0;
  const int* iter = nullptr;
  double* alpha = nullptr;
  double* sse_out = nullptr;
};
void launch_noise_draw(const NoiseDrawArgs& a, hipStream_t s);

struct EvalArgs {
  int ntest = 0, nmodes = 2, K = 0, KP = 0;
  const int* idx = nullptr;      // ntest x nmodes
  const double* vals = nullptr;  // ntest
  const float* f[kMaxOther + 1] = {};
  double mean = 0;
  double* pred = nullptr;
  double* pred_var = nullptr;
  double* partial = nullptr;     // per-block (se, se_avg)
  int nblk = 0;
  const int* iter = nullptr;
  int burnin = 0;
  double* out = nullptr;         // (rmse1, rmse_avg)
  // probit (ProbitNoise::evalModel, noisemodels.cpp:129-159): predictions
  // Phi(u.v) without the mean; pred1 receives the one-sample predictions
  int probit = 0;
  double* pred1 = nullptr;
};
int eval_blocks(int ntest);
void launch_eval(const EvalArgs& a, hipStream_t s);

void launch_advance_iter(int* iter, hipStream_t s);

// --- probit augmentation (latentprior.cpp:387-414) ---
struct ProbitZArgs {
  long long nnz = 0;
  int KP = 0;
  const int* row_of = nullptr;  // nnz: row of each observation (mode order)
  const int* idx = nullptr;     // nnz: other-mode index
  const double* y = nullptr;    // nnz: raw values
  const float* U = nullptr;     // this mode's factor, KP x N
  const float* V = nullptr;     // other mode's factor
  float* z = nullptr;           // nnz out
  uint64_t seed = 0;
  const int* iter = nullptr;
  int mode = 0;
};
void launch_probit_z(const ProbitZArgs& a, hipStream_t s);
void launch_truncnorm_probe(int n, const double* mean, const double* p, uint64_t seed, uint32_t c0, uint32_t it,
                            uint32_t tag, double* out, double* inv, hipStream_t s);

// --- RNG probes (parity tests) ---
void launch_rng_probe(uint64_t seed, uint32_t c0, uint32_t it, uint32_t tag, int n, uint32_t* bits, float* z32,
                      double* z64, hipStream_t s);
void launch_gamma_probe(uint64_t seed, const double* shapes, int n, uint32_t it, uint32_t tag, double* out,
                        hipStream_t s);

// --- sparse linear algebra for the link matrix (FP64) ---
struct CsrDev {
  int nrow = 0, ncol = 0;
  const int* row_ptr = nullptr;
  const int* cols = nullptr;
  const double* vals = nullptr;  // null = binary
  // optional nnz-balanced work list (used by launch_spmm when nseg > 0):
  // segments (row, first nnz, end nnz, slot) of at most kSpmmSegNnz
  // nonzeros; slot < 0: the segment is its whole row and writes the output,
  // else it writes partial[slot] and the rows split into several segments
  // (multi_row, slots multi_slot0[m] .. multi_slot0[m + 1] - 1) are summed
  // in slot order afterwards
  int nseg = 0;
  // the first nlong segments (the split rows' segments, then the whole rows
  // of >= kSpmmLongNnz nonzeros, longest first) take a wave each; the rest
  // (whole rows, by decreasing length) go four to a wave, except the last
  // nempty (rows without nonzeros), sixteen to a wave
  int nlong = 0, nempty = 0;
  const int4* seg = nullptr;
  int nmulti = 0;
  const int* multi_row = nullptr;
  const int* multi_slot0 = nullptr;
  double* partial = nullptr;  // slots x nr
  // optional dense block of a binary F (launch_spmm with nr = 32 only): the
  // ndp (a multiple of 16, at most kDenseMaxCols) densest columns of F in bit
  // form; dense_role 1 on F (rows), 2 on its transpose (the dense columns
  // are rows there)
  int dense_role = 0;
  int ndp = 0, nd = 0;        // padded and real count of dense columns
  int nfrow = 0;              // rows of F
  int ndw = 0;                // words per row of row_bits: ndp / 32 rounded up
  int nrw = 0;                // words per dense column of col_bits: (nfrow + 31) / 32
  const int* dcol = nullptr;  // ndp column ids of F (padding: column 0, no bits)
  const uint32_t* row_bits = nullptr;  // role 1: bit j of row r set when F[r][dcol[j]] = 1
  const int* rem_ptr = nullptr;        // role 1: F without the dense columns (binary CSR)
  const int* rem_cols = nullptr;
  const int* rem_fix = nullptr;        // role 1: the first kDenseRem of rem_cols per row (padding: column 0)
  const uint32_t* col_bits = nullptr;  // role 2: bit r of dense column j at word j * nrw + r / 32
  int nkc = 0;                // role 2: chunks of kDenseRowChunk rows of F per dense column
  int dslot0 = 0;             // role 2: partial slot of (column j, chunk k) = dslot0 + j * nkc + k
  // role 2: the first row at or after r with a nonzero (nrow past the last);
  // rows without nonzeros and none within kEmptyReach rows either way get
  // no work unit: their output (reg * add) is formed where it is read
  const int* next_nonempty = nullptr;
};
constexpr int kEmptyReach = 31;
constexpr int kDenseMaxCols = 288;
constexpr int kDenseRowChunk = 256;
constexpr int kDenseRem = 8;

// --- F^T F branch of MacauPrior::sample_beta (latentprior.cpp:280-289) ---
// fail-flag value set when F^T F + lambda_beta I is not positive definite
constexpr int kFailDenseChol = 1 << 29;
// the Normal-Wishart draw's Cholesky of its scale or precision matrix failed
constexpr int kFailNormalWishart = 1 << 30;
// F^T F (ncol x ncol, col-major, lower triangle; the upper is zeroed) of the
// CSR F and its transpose Ft, in the reference's accumulation order
void launch_ftf(const CsrDev& F, const CsrDev& Ft, double* FtF, hipStream_t s);
// X (n x nrhs row-major, holding B on entry) <- (FtF + reg I)^{-1} X through
// a blocked Cholesky in the n x n work matrix W; *reg is a device scalar;
// a non-positive pivot sets *fail to kFailDenseChol
void launch_ftf_solve(const double* FtF, int n, const double* reg, double* W, double* X, int nrhs, int* fail,
                      hipStream_t s);
// the same solve through F^T F = Q diag(d) Q^T: spectral_factor overwrites
// A (F^T F, lower triangle) with Q, fills d (n) and Qt = Q^T and returns
// false with a reason when the eigensolver is unavailable or fails;
// launch_spectral_solve sets X = Q diag(d + *reg)^{-1} Q^T B through
// Y (n x nrhs) and part (plan.S x n x nrhs), a non-positive d_j + *reg
// setting *fail to kFailDenseChol
struct SpectralPlan {
  int S = 1;       // inner-dimension chunks
  int tchunk = 0;  // chunk depth
};
SpectralPlan spectral_plan(int n, int nrhs);
bool spectral_factor(double* A, int n, double* d, double* Qt, hipStream_t s, std::string* why);
void launch_spectral_solve(const double* Q, const double* Qt, const double* d, int n, const double* reg,
                           const double* B, double* Y, double* part, const SpectralPlan& p, double* X, int nrhs,
                           int* fail, hipStream_t s);
constexpr int kSpmmSegNnz = 256;
constexpr int kSpmmLongNnz = 32;
// out[:, r] = sum_c A[r,c] in[:, c] (+ (*reg) * add[:, r]); vectors have ld = nr
// (any nr; beyond 64 the columns are taken in further passes of 64);
// reg is a device scalar, read only when add is non-null; nothing is done
// when skip is non-null and *skip != 0 (a converged block CG)
void launch_spmm(const CsrDev& A, const double* in, int nr, double* out, const double* reg, const double* add,
                 hipStream_t s, const double* skip = nullptr);
// partial[0][q] = sum_b partial[b][q] (fixed order), q < width
void launch_partial_reduce(double* partial, int nblk, int width, hipStream_t s);
// copy rows [r0, r0 + nr) of a (K x n, ld K) matrix to (nr x n, ld nr), or back
void launch_rows_copy(const double* src, int lds, double* dst, int ldd, int r0s, int r0d, int nr, int n, hipStream_t s);
// sum of squares of n values (partial per block, FP64)
void launch_sumsq_f32(const float* x, size_t n, double* partial, int nblk, hipStream_t s);
void launch_sumsq_f64(const double* x, size_t n, double* partial, int nblk, hipStream_t s);
// sum_f a[:,f] b[:,f]^T (nr x nr, col-major) accumulated per block into
// partial[nblk][nr * nr] and reduced in place: the total is in partial[0]
// when the launch completes (nblk = dots_blocks(n))
int dots_blocks(int n);
void launch_dots(const double* a, const double* b, int nr, int n, double* partial, int nblk, hipStream_t s,
                 const double* skip = nullptr, bool reduce = true, const int* nxt = nullptr,
                 const double* reg = nullptr);

struct CgState {
  int nr = 0, n = 0;
  double* X = nullptr;
  double* R = nullptr;
  double* P = nullptr;
  double* KPm = nullptr;
  double* inner = nullptr;
  double* partial = nullptr;
  int nblk = 0;
  double* small = nullptr;  // RtR, RtR2, PtKP, A, Psi, norms, flags
  // nr = 32 with a dense-block plan: the next_nonempty list of F^T and the
  // regularisation; a group of at most 32 features with no nonzero in F^T
  // takes KP = reg * P where it is read (KP is not written there)
  const int* nxt = nullptr;
  const double* reg = nullptr;
};
void launch_cg_init(const CgState& c, const double* B, hipStream_t s);
void launch_cg_step_a(const CgState& c, hipStream_t s);   // A from PtKP partials
void launch_cg_update_xr(const CgState& c, hipStream_t s);
void launch_cg_step_b(const CgState& c, double tolsq, hipStream_t s);  // convergence + Psi
void launch_cg_update_p(const CgState& c, hipStream_t s);
void launch_cg_finish(const CgState& c, hipStream_t s);

struct BetaArgs {
  int K = 0, KP = 0, N = 0, nfeat = 0;
  const float* U = nullptr;
  const double* mu = nullptr;
  const double* Lchol = nullptr;
  const double* lambda_beta = nullptr;
  uint64_t seed = 0;
  const int* iter = nullptr;
  int mode = 0;
};
void launch_beta_tmp(const BetaArgs& a, double* tmp, hipStream_t s);
void launch_beta_noise_add(const BetaArgs& a, double* Fty, hipStream_t s);
// tmp: K x nrow FP64 scratch
void launch_uhat_finish(const CsrDev& F, const double* beta, int K, int KP, const double* mu, const double* Lambda,
                        float* Uhat, float* Lmu_rows, double* tmp, hipStream_t s);
void launch_lambda_beta(int K, double nfeat, const double* bbt_partial, int nblk, const double* Lambda,
                        double* lambda_beta, uint64_t seed, const int* iter, int mode, hipStream_t s);

// --- univariate prior MacauOnePrior (macauoneprior.cpp:189-289) ---
struct BetaUniArgs {
  int K = 0, KP = 0, N = 0, nfeat = 0;
  const float* U = nullptr;          // KP x N
  const double* uhat = nullptr;      // K x N (previous Uhat, FP64)
  const double* mu = nullptr;        // K
  const double* Lambda = nullptr;    // K x K, diagonal
  double* lambda_beta = nullptr;     // K (read by the sweep, written by launch_lambda_beta_uni)
  const double* colsq = nullptr;     // nfeat: sum of squares of each column of F
  const int* Ft_rowptr = nullptr;    // F^T as CSR (nfeat rows), no repeated entries in a row
  const int* Ft_cols = nullptr;
  const double* Ft_vals = nullptr;   // values (1 for binary side information), never null
  double* beta = nullptr;            // K x nfeat
  double* ainv = nullptr;            // K x nfeat scratch
  double* nz = nullptr;              // K x nfeat scratch
  double* Zw = nullptr;              // K x N scratch (used when N exceeds the LDS rows)
  // the features with non-empty columns in order, in batches of 4 records
  // {feature (-1 = unused slot), column begin, column end, wide}: up to four
  // consecutive columns of at most 16 entries with pairwise disjoint rows
  // (wide = 0), or one longer column repeated in all four records (wide = 1)
  const int4* batch_meta = nullptr;
  int nbatch = 0;
  uint64_t seed = 0;
  const int* iter = nullptr;
  int mode = 0;
};
// Lambda <- diag(Lambda), Lambda_f <- its diagonal, Lmu_f = lambda mu
void launch_uni_diag(int K, int KP, double* Lambda, float* Lambda_f, float* Lmu_f, const double* mu, hipStream_t s);
// sample_beta: the per-latent feature sweeps
void launch_beta_uni(const BetaUniArgs& a, hipStream_t s);
// partial: lambda_beta_uni_partials(K) doubles of scratch
int lambda_beta_uni_partials(int K);
void launch_lambda_beta_uni(const BetaUniArgs& a, double* partial, hipStream_t s);
bool beta_uni_in_lds(int N);

}  // namespace mgpu
