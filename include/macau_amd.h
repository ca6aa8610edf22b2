/*
 * macau_amd.h -- C ABI of the MI355X-native Gibbs sampler for BPMF / Macau.
 *
 * This is the drop-in boundary for the reference's sampling step.  The
 * reference binds its C++ core from Cython (python/macau/macau.pxd:63-84);
 * every entry point below replaces one of those bindings or one hot-path
 * function of the C++ core, cited next to it (file:line in jaak-s/macau).
 *
 * Conventions
 *   - plain C types only; arrays are host pointers, copied in and out (the
 *     library never retains a caller pointer after the call returns);
 *   - matrices are column-major like Eigen::MatrixXd (a K x N factor matrix
 *     stores latent vector n as K contiguous doubles);
 *   - functions returning int give 0 on success and -1 on error; the message
 *     of the last error on the calling thread is macau_last_error();
 *     reference exceptions keep their text (e.g. "Cholesky Decomposition
 *     failed!", latentprior.cpp:375; "Less 2 priors are not supported.",
 *     macau.cpp:75);
 *   - one host thread per model; a model owns one HIP stream on device 0.
 */
#ifndef MACAU_AMD_H
#define MACAU_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct macau_model macau_model;

/* sampler kinds (macau_set_sampler) */
#define MACAU_SAMPLER_PLAIN_FP32 0 /* one work-group per row, FP32 Gram       */
#define MACAU_SAMPLER_BATCHED    1 /* nnz-bucketed, split-BF16 MFMA Gram      */
/* macau_sample_latents only: the batched sampler without its refinement step */
#define MACAU_SAMPLER_BATCHED_NOREFINE 2

const char* macau_last_error(void);
const char* macau_version(void);

/* ---- factories: make_macau_fixed / make_macau_adaptive / make_macau_probit
 *      (macau.h:83-85, macau.cpp:207-228; macau.pxd:82-84) ---- */
macau_model* macau_make_fixed(int nmodes, int num_latent, double precision);
macau_model* macau_make_adaptive(int nmodes, int num_latent, double sn_init, double sn_max);
/* ProbitNoise (noisemodels.h:88-103): matrix data (nmodes = 2); with more
 * modes macau_init / macau_run fail like the reference's sampler and
 * evaluation (latentprior.cpp:112-116, noisemodels.cpp:177-181) */
macau_model* macau_make_probit(int nmodes, int num_latent);
void macau_free(macau_model* m); /* replaces `del macau` (macau.pyx:449) */

/* ---- priors: BPMFPrior(num_latent) (latentprior.h:53-79, macau.pxd:45-49)
 *      and MacauPrior<SparseDoubleFeat> (latentprior.h:82-125,
 *      macau.pyx:186-192, make_prior macau.pyx:154-192; side information as COO, rows = data mode) ---- */
int macau_add_bpmf_prior(macau_model* m);
int macau_add_macau_prior(macau_model* m, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                          const double* vals /* NULL = binary */, double lambda_beta, double tol);
/* make_prior(side, num_latent, max_ff_size, lambda_beta, tol) (macau.pyx:154-192):
 * the same prior, and the link matrix is solved through the Cholesky factor of
 * F^T F + lambda_beta I (compute_ff, latentprior.cpp:199-203, 280-289) when
 * ncol <= max_ff_size, by block CG otherwise; macau_add_macau_prior always
 * uses the block CG.  The Python layer passes max_ff_size = 10000 as
 * setSidePriors does (macau.pyx:334). */
int macau_add_macau_prior_ff(macau_model* m, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                             const double* vals /* NULL = binary */, double lambda_beta, double tol,
                             int max_ff_size);
/* make_one_prior(side, num_latent, lambda_beta) (macau.pyx:194-214):
 * MacauOnePrior<SparseFeat / SparseDoubleFeat> (macauoneprior.h:8-51) with
 * setLambdaBeta(lambda_beta); the univariate=True path of setSidePriors
 * (macau.pyx:328-333).  Not available with probit noise. */
int macau_add_macau_one_prior(macau_model* m, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                              const double* vals /* NULL = binary */, double lambda_beta);

/* ---- Macau API (macau.h:35-53; macau.pxd:63-84) ---- */
int macau_set_samples(macau_model* m, int burnin, int nsamples);
/* idx: nnz x nmodes int32, column-major (toMatrix, sparsetensor.cpp:203-211) */
int macau_set_relation_data(macau_model* m, const int* idx, int nmodes, const double* values, int nnz,
                            const int* dims);
int macau_set_relation_data_test(macau_model* m, const int* idx, int nmodes, const double* values, int nnz,
                                 const int* dims);
int macau_set_verbose(macau_model* m, int verbose);
int macau_set_save_model(macau_model* m, int save);
int macau_set_save_prefix(macau_model* m, const char* prefix);
int macau_run(macau_model* m); /* MacauX::run (macau.cpp:94-148) */
double macau_get_rmse_test(macau_model* m);
long long macau_get_test_nonzeros(macau_model* m);
int macau_get_predictions(macau_model* m, double* out); /* ntest */
int macau_get_stds(macau_model* m, double* out);        /* ntest (macau.cpp:175-188) */
int macau_get_test_data(macau_model* m, double* out);   /* ntest x (nmodes+1), col-major */

/* ---- extensions (no reference counterpart) ---- */
int macau_set_seed(macau_model* m, unsigned long long seed);   /* counter-based RNG key */
int macau_set_sampler(macau_model* m, int kind);               /* MACAU_SAMPLER_* */
int macau_sampler_supported(int kind, int num_latent, int nmodes); /* 1 if available */
int macau_set_graph(macau_model* m, int on);                   /* replay iterations as a HIP graph */
int macau_init(macau_model* m);                                /* MacauX::init (macau.cpp:72-86) */
int macau_iterate(macau_model* m, int n);                      /* n Gibbs iterations, asynchronous */
int macau_synchronize(macau_model* m);                         /* wait; raise sampler failures */
int macau_get_iteration(macau_model* m);
/* n iterations without graph replay, timed with HIP events on the model's
 * stream; phase_ms[2 * nmodes + 2] receives the mean ms of: sampling of each
 * mode, the prior update of each mode, the noise update, the evaluation */
int macau_profile(macau_model* m, int n, double* phase_ms);
/* RMSE of the one-sample / averaged predictions; for probit models the
 * predictions are probabilities and these are root Brier scores */
double macau_get_rmse_onesample(macau_model* m);
double macau_get_rmse_avg(macau_model* m);
/* probit models: test AUC of the one-sample and averaged predictions
 * (ProbitNoise::auc_test_onesample / auc_test, noisemodels.cpp:156-157) */
int macau_get_auc(macau_model* m, double* auc_onesample, double* auc_avg);
double macau_get_noise_precision(macau_model* m);
double macau_get_mean_value(macau_model* m);
int macau_get_dims(macau_model* m, int* dims);
/* current samples / hyperparameters of one mode, FP64 */
int macau_get_latents(macau_model* m, int mode, double* out);              /* K x N_mode */
int macau_get_hyper(macau_model* m, int mode, double* mu, double* Lambda); /* K, K x K */
int macau_get_link(macau_model* m, int mode, double* beta);                /* K x nfeat */
double macau_get_link_lambda(macau_model* m, int mode);
int macau_get_last_cg_iters(macau_model* m, int mode);
/* Which solve the F^T F branch uses after macau_setup (no counterpart in the
 * reference, which always factors F^T F + lambda_beta I by Cholesky,
 * latentprior.cpp:280-288): 0 = the mode has no F^T F branch, 1 = blocked
 * Cholesky every iteration, 2 = eigendecomposition of F^T F at setup and two
 * products per iteration (the default; environment variable
 * MACAU_AMD_FTF_SOLVER=cholesky selects 1). */
int macau_get_ftf_solver(macau_model* m, int mode);

/* ---- refinement of ill-conditioned rows (batched sampler; no reference
 *      counterpart: it guards the reference's FP64 solve,
 *      latentprior.cpp:373-384, against the FP32 / split-BF16 arithmetic) ----
 * Beside each row's Cholesky the batched sampler estimates the row's
 * conditioning with three probe solves y = L^{-1} D^{1/2} x (D = diag P;
 * x = the row's normals, ones, alternating signs) done with the forward
 * solve: est = K max_x |y|^2 / |x|^2, an estimate of the condition number
 * of the Jacobi-scaled D^{-1/2} P D^{-1/2} (K for a diagonal P); a row
 * above `threshold` (default 500) with at least max(num_latent, 128)
 * nonzeros (fewer: not probed) gets one step
 * of iterative refinement against the exact precision matrix (residual from
 * per-observation residuals, FP64 sums).  macau_set_refine (before
 * macau_init) switches it off (on = 0) or sets the threshold (> 0);
 * macau_get_refine_stats reports the rows flagged and refined so far (a
 * half-sweep refines at most 4096 rows).  num_latent above 64 is sampled
 * without probes or refinement. */
int macau_set_refine(macau_model* m, int on, double threshold);
int macau_get_refine_stats(macau_model* m, long long* flagged, long long* refined);

/* ---- autotuner of the batched sampler (no reference counterpart: the
 *      reference schedules every row loop as OpenMP "dynamic, 2",
 *      latentprior.cpp:73) ----
 * Knobs per mode: heavy-row segment length, and the work-group caps of the
 * light-row, heavy-row and segment kernels.  macau_autotune (after
 * macau_init, before or between iterations) times every candidate shape of
 * every mode on the device with `reps` half-sweeps each, keeps the fastest
 * (the default unless a candidate is more than 1% faster), leaves the chain's
 * state as it was, writes each mode's half-sweep ms with the default and the
 * chosen shape to default_ms[nmodes] / tuned_ms[nmodes] (either may be
 * NULL) and, when record_path is not NULL, appends the choices to that
 * table.  macau_init looks the shape up by (GPU architecture, padded latent
 * size, other modes, rows, nonzeros, log2 row-degree histogram) in the table
 * named by the environment variable MACAU_AMD_TUNE_CACHE, then in
 * tuning.txt beside the library.  Plain-sampler models have nothing to tune.
 * macau_get_tuning writes (segment length, light cap, heavy cap, segment
 * cap) of a mode to knobs[4] and returns 1 if they come from the tuner or
 * the table, 0 for the defaults, -1 on error. */
int macau_autotune(macau_model* m, int reps, const char* record_path, double* default_ms, double* tuned_ms);
int macau_get_tuning(macau_model* m, int mode, int* knobs);

/* ---- batches: many independent models as one launch set (no reference
 *      counterpart: the reference runs one model per process,
 *      macau.cpp:94-148) ----
 * A batch advances N initialised models (any sizes, latent dimensions,
 * priors and noise models; no Macau prior solved by block CG) by one Gibbs
 * iteration per replay of one HIP graph whose branches are the models'
 * iterations.  Each model keeps its own state, draws and iteration counter:
 * the getters above read it as usual after macau_batch_synchronize.  The
 * batch does not own the models; free the batch before them. */
typedef struct macau_batch macau_batch;
macau_batch* macau_batch_create(macau_model* const* models, int n);
void macau_batch_free(macau_batch* b);
/* 1 (default): one graph per iteration of all models; 0: each model replays
 * its own iteration graph on its own streams, all queued before any wait */
int macau_batch_set_graph(macau_batch* b, int on);
int macau_batch_iterate(macau_batch* b, int n);     /* n iterations of every model, asynchronous */
int macau_batch_synchronize(macau_batch* b);        /* wait; raise sampler failures */

/* ---- stateless hot-path entry points ---- */

/* One half-sweep: ILatentPrior::sample_latents(U, mat, mean, samples, alpha, K)
 * (latentprior.h:19-20) for BPMFPrior (latentprior.cpp:69-77) and
 * MacauPrior (mu_rows != NULL, latentprior.cpp:218-227), and the tensor form
 * sample_latent_tensor (latentprior.cpp:118-161) when nother > 1.
 *   rows: nrows + 1 offsets into idx (nnz x nother, row-major) and vals;
 *   others[t]: K x other_n[t] column-major factors;
 *   draws: normal k of row n is stream (c0 = n, iter, tag = (1 << 24) | mode). */
int macau_sample_latents(int num_latent, int nother, int nrows, const long long* row_ptr, const int* idx,
                         const double* vals, double mean, const double* const* others, const int* other_n,
                         double alpha, const double* mu, const double* Lambda, const double* mu_rows,
                         unsigned long long seed, int iter, int mode, int sampler, double* U_out);

/* CondNormalWishart (mvnormal.cpp:314-336): U is K x N column-major. */
int macau_cond_normal_wishart(int num_latent, int N, const double* U, const double* mu0, double kappa,
                              const double* T, double nu, unsigned long long seed, int iter, int mode,
                              double* mu_out, double* Lambda_out);

/* Allocation-free CondNormalWishart.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same draw as
 * macau_cond_normal_wishart, but every array is a DEVICE pointer owned by the
 * caller (d_U: K x N column-major FP64; d_mu0: K; d_T: K x K; outputs d_mu_out:
 * K, d_Lambda_out: K x K, *d_status: 0, or nonzero when a matrix was not
 * positive definite), the work is queued asynchronously on `stream`
 * (a hipStream_t; NULL = the null stream) and no device memory is allocated
 * or freed.  Working memory is one static device array of 2,146,560 bytes per
 * GPU (held until the process ends), so num_latent must be in [1, 64]; larger
 * values fail with an error.  Calls that overlap on one GPU share that array
 * and race. */
int macau_cond_normal_wishart_noalloc(int num_latent, int N, const double* d_U, const double* d_mu0, double kappa,
                                      const double* d_T, double nu, unsigned long long seed, int iter, int mode,
                                      double* d_mu_out, double* d_Lambda_out, int* d_status, void* stream);

/* solve_blockcg (linop.h:228-343; Python blockcg, macau.pyx:67-91): solves
 * X (F^T F + reg I) = B for nrhs <= 40 right-hand sides; B and X are
 * nrhs x nfeat column-major; F is COO nrow x ncol (vals NULL = binary).
 * *iters receives the iteration count. */
int macau_solve_blockcg(int nrhs, int nrow, int ncol, long long nnz, const int* rows, const int* cols,
                        const double* vals, double reg, const double* B, double tol, double* X, int* iters);

/* RNG probes: Philox words, FP32 and FP64 normals of stream (c0, iter, tag)
 * for j = 0..n-1, and FP64 gamma draws with shapes[j] at c0 = j. */
int macau_rng_probe(unsigned long long seed, unsigned c0, unsigned iter, unsigned tag, int n, unsigned* bits,
                    float* z32, double* z64);
int macau_gamma_probe(unsigned long long seed, const double* shapes, int n, unsigned iter, unsigned tag,
                      double* out);
/* out[j] = rand_truncnorm(mean[j], 1, 0) (truncnorm.cpp:9-37) keyed by
 * (c0 + j, iter, tag); inv[j] = inv_norm_cdf(p[j]) */
int macau_truncnorm_probe(unsigned long long seed, unsigned c0, unsigned iter, unsigned tag, int n,
                          const double* mean, const double* p, double* out, double* inv);

#ifdef __cplusplus
}
#endif
#endif
