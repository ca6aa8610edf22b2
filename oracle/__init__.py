"""FP64 CPU oracle for the jaak-s/macau Gibbs sampling step.

TEST INFRASTRUCTURE ONLY.  tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, as the checker
and the reported CPU baseline.  The product (``macau_amd``) never imports it.

The numeric core is ``oracle.c`` (built into ``liboracle.so`` by
``oracle/Makefile``).  This module adds the data layer restated in numpy:

* ``matrix_rows`` -- ``MatrixData::setTrain`` (sparsetensor.cpp:92-106)
  through ``sparseFromIJV`` (bpmfutils.h:57-65): Eigen ``setFromTriplets``
  sums duplicates and sorts inner indices; mode 0 iterates ``Yt`` columns
  (rows of Y), mode 1 iterates ``Y`` columns (latentprior.cpp:30-46).
* ``sparse_mode`` -- ``SparseMode::init`` (sparsetensor.cpp:11-60): stable
  grouping by one mode, the other modes' indices kept in order.
* ``matrix_test`` / ``tensor_test`` -- the test-cell order of
  ``getTestData`` (sparsetensor.cpp:63-76 CSC order; 178-191 mode-0 order).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

TAG_LATENT, TAG_WISHART_Z, TAG_WISHART_CHI, TAG_NW_MU = 1, 2, 3, 4
TAG_NOISE, TAG_BETA_U, TAG_BETA_F, TAG_LAMBDA_BETA = 5, 6, 7, 8
TAG_PROBIT = 9
TAG_BETA_ONE = 10
MAX_MODES = 10  # ORC_MAX_MODES


def tag(stream, mode=0):
    return (stream << 24) | mode


_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_dbl = ctypes.c_double


class OrcRows(ctypes.Structure):
    _fields_ = [("nrows", _i64), ("nother", ctypes.c_int), ("row_ptr", _p), ("idx", _p), ("vals", _p)]


class OrcCsr(ctypes.Structure):
    _fields_ = [("nrow", _i64), ("ncol", _i64), ("row_ptr", _p), ("cols", _p), ("vals", _p)]


class OrcRunCfg(ctypes.Structure):
    _fields_ = [
        ("nmodes", ctypes.c_int), ("K", ctypes.c_int), ("is_tensor", ctypes.c_int),
        ("dims", _i64 * MAX_MODES), ("train", OrcRows * MAX_MODES), ("mean", _dbl), ("nnz", _i64),
        ("ntest", _i64), ("test_idx", _p), ("test_vals", _p),
        ("noise_adaptive", ctypes.c_int), ("precision", _dbl), ("sn_init", _dbl), ("sn_max", _dbl),
        ("macau", ctypes.c_int * MAX_MODES), ("F", OrcCsr * MAX_MODES), ("Ft", OrcCsr * MAX_MODES),
        ("lambda_beta0", _dbl), ("tol", _dbl), ("burnin", ctypes.c_int), ("nsamples", ctypes.c_int),
        ("seed", _u64), ("max_iters", ctypes.c_int),
        ("pred", _p), ("pred_var", _p), ("U_out", _p * MAX_MODES), ("rmse", _dbl), ("rmse1", _dbl),
        ("alpha", _dbl), ("alpha_trace", _p), ("rmse_trace", _p), ("lambda_beta_out", _p),
        ("cg_iters_last", ctypes.c_int), ("use_ftf", ctypes.c_int * MAX_MODES),
        ("noise_probit", ctypes.c_int), ("auc", _dbl), ("auc1", _dbl), ("uni", ctypes.c_int * MAX_MODES),
        ("mu_out", _p * MAX_MODES), ("Lambda_out", _p * MAX_MODES),
    ]


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        _LIB = _load(path)
    return _LIB


class use_library:
    """Context manager: route every oracle call through the library at
    ``path`` (a variant built from oracle.c with a planted mistake, used by
    the tests that show the pins catch such mistakes)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _LIB
        self.saved = lib()
        _LIB = _load(self.path)
        return _LIB

    def __exit__(self, *exc):
        global _LIB
        _LIB = self.saved
        return False


def _load(path):
    """Load an oracle library and declare its entry points."""
    L = ctypes.CDLL(path)
    L.orc_philox4x32_10.argtypes = [_p, _p, _p]
    L.orc_normal.restype = _dbl
    L.orc_normal.argtypes = [_u64, _u32, _u32, _u32, _u32]
    L.orc_normals.argtypes = [_u64, _u32, _u32, _u32, _u32, _p]
    L.orc_gamma.restype = _dbl
    L.orc_gamma.argtypes = [_u64, _dbl, _u32, _u32, _u32]
    L.orc_sample_latents.restype = _i64
    L.orc_sample_latents.argtypes = [ctypes.c_int, ctypes.POINTER(OrcRows), _dbl, _p, _dbl, _p, _p, _p,
                                     _u64, _u32, ctypes.c_int, _p]
    L.orc_cond_normal_wishart.argtypes = [ctypes.c_int, _i64, _p, _p, _dbl, _p, _dbl, _u64, _u32,
                                          ctypes.c_int, _p, _p]
    L.orc_normal_wishart_from_stats.argtypes = [ctypes.c_int, _i64, _p, _p, _p, _dbl, _p, _dbl, _u64,
                                                _u32, ctypes.c_int, _p, _p]
    L.orc_train_sse.restype = _dbl
    L.orc_train_sse.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcRows), _dbl, _p]
    L.orc_noise_alpha.restype = _dbl
    L.orc_noise_alpha.argtypes = [_dbl, _i64, _dbl, ctypes.c_int, _dbl, _u64, _u32]
    L.orc_eval_rmse.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _p, _p, _dbl, _p, ctypes.c_int, _p,
                                _p, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]
    L.orc_csr_A_mul_Bn.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCsr), _p, _p]
    L.orc_AtA_mul_B.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCsr), ctypes.POINTER(OrcCsr), _dbl, _p, _p]
    L.orc_blockcg.restype = ctypes.c_int
    L.orc_blockcg.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCsr), ctypes.POINTER(OrcCsr), _dbl, _p, _dbl, _p]
    L.orc_blockcg_blocked.restype = ctypes.c_int
    L.orc_blockcg_blocked.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCsr), ctypes.POINTER(OrcCsr), _dbl, _p,
                                      _dbl, ctypes.c_int, ctypes.c_int, _p]
    L.orc_posterior_lambda_beta.argtypes = [ctypes.c_int, _i64, _p, _p, _dbl, _dbl, ctypes.POINTER(_dbl),
                                            ctypes.POINTER(_dbl)]
    L.orc_at_mul_a.argtypes = [ctypes.POINTER(OrcCsr), ctypes.POINTER(OrcCsr), _p]
    L.orc_ftf_solve.restype = ctypes.c_int
    L.orc_ftf_solve.argtypes = [ctypes.c_int, _i64, _p, _dbl, _p, _p]
    L.orc_norm_cdf.restype = _dbl
    L.orc_norm_cdf.argtypes = [_dbl]
    L.orc_inv_norm_cdf.restype = _dbl
    L.orc_inv_norm_cdf.argtypes = [_dbl]
    L.orc_truncnorm0.restype = _dbl
    L.orc_truncnorm0.argtypes = [_dbl, _u64, _u32, _u32, _u32]
    L.orc_probit_z.argtypes = [ctypes.c_int, ctypes.POINTER(OrcRows), _p, _p, _u64, _u32, ctypes.c_int, _p]
    L.orc_auc.restype = _dbl
    L.orc_auc.argtypes = [_i64, _p, _p]
    L.orc_eval_probit.argtypes = [ctypes.c_int, _i64, _p, _p, _p, ctypes.c_int, _p, _p, _p,
                                  ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]
    L.orc_sample_latents_uni.argtypes = [ctypes.c_int, ctypes.POINTER(OrcRows), _dbl, _p, _dbl, _p, _p, _u64,
                                         _u32, ctypes.c_int, _p]
    L.orc_at_mul_bt_col.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCsr), _i64, _p, _p]
    L.orc_add_acol_mul_bt.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCsr), _i64, _p, _p]
    L.orc_beta_sweep_uni.argtypes = [ctypes.c_int, _i64, ctypes.POINTER(OrcCsr), _p, _p, _p, _p, _p, _p, _p,
                                     _u64, _u32, ctypes.c_int]
    L.orc_lambda_beta_uni.argtypes = [ctypes.c_int, _i64, _p, _u64, _u32, ctypes.c_int, _p]
    L.orc_macau_one_update_prior.restype = ctypes.c_int
    L.orc_macau_one_update_prior.argtypes = [ctypes.c_int, ctypes.c_int, _p, _i64, ctypes.POINTER(OrcCsr),
                                             ctypes.POINTER(OrcCsr), _p, _p, _p, _p, _p, _p, _u64, _u32]
    L.orc_run.restype = ctypes.c_int
    L.orc_run.argtypes = [ctypes.POINTER(OrcRunCfg)]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_get_threads.restype = ctypes.c_int
    L.orc_sample_beta.restype = ctypes.c_int
    L.orc_sample_beta.argtypes = [ctypes.c_int, ctypes.c_int, _p, _i64, _p, _p, ctypes.POINTER(OrcCsr),
                                  ctypes.POINTER(OrcCsr), _p, _dbl, _dbl, _u64, _u32, _p,
                                  ctypes.POINTER(ctypes.c_int)]
    L.orc_macau_update_prior.restype = ctypes.c_int
    L.orc_macau_update_prior.argtypes = [ctypes.c_int, ctypes.c_int, _p, _i64, ctypes.POINTER(OrcCsr),
                                         ctypes.POINTER(OrcCsr), _p, _p, ctypes.c_int, _p, _p,
                                         ctypes.POINTER(_dbl), _dbl, _p, _p, _u64, _u32,
                                         ctypes.POINTER(ctypes.c_int)]
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------------------------------------------------------- RNG
def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def normals(seed, c0, n, it, tg):
    out = np.zeros(n, dtype=np.float64)
    lib().orc_normals(seed, c0, n, it, tg, _ptr(out))
    return out


def gamma(seed, shape, c0, it, tg):
    return lib().orc_gamma(seed, shape, c0, it, tg)


# ---------------------------------------------------------------- data layer
class Rows:
    """Sparse rows of one mode: row_ptr (int64), idx (nnz x nother int32), vals (float64)."""

    def __init__(self, row_ptr, idx, vals):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(idx, dtype=np.int32)
        if self.idx.ndim == 1:
            self.idx = self.idx.reshape(-1, 1)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)

    @property
    def nrows(self):
        return len(self.row_ptr) - 1

    def c(self):
        return OrcRows(self.nrows, self.idx.shape[1], _ptr(self.row_ptr), _ptr(self.idx), _ptr(self.vals))


def sum_duplicates(rows, cols, vals, nrows, ncols):
    """setFromTriplets: duplicates summed; returns (row, col, val) in CSC order."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    key = cols * nrows + rows
    order = np.argsort(key, kind="stable")
    ks = key[order]
    if len(ks) == 0:
        return rows[:0], cols[:0], vals[:0]
    starts = np.concatenate([[0], np.nonzero(np.diff(ks))[0] + 1])
    v = np.add.reduceat(vals[order], starts)
    uk = ks[starts]
    return (uk % nrows).astype(np.int64), (uk // nrows).astype(np.int64), v


def _group(keys, nkeys, other, vals):
    order = np.argsort(keys, kind="stable")
    counts = np.bincount(keys, minlength=nkeys)
    row_ptr = np.zeros(nkeys + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    return Rows(row_ptr, other[order], vals[order])


def matrix_rows(rows, cols, vals, nrows, ncols):
    """MatrixData::setTrain: returns ([mode0 Rows, mode1 Rows], mean, nnz)."""
    r, c, v = sum_duplicates(rows, cols, vals, nrows, ncols)
    mean = v.sum() / len(v) if len(v) else np.nan
    # mode 0: rows of Y with column indices ascending (CSC of Yt)
    o0 = np.lexsort((c, r))
    m0 = _group(r[o0], nrows, c[o0].reshape(-1, 1), v[o0])
    # mode 1: columns of Y with row indices ascending (CSC of Y)
    m1 = _group(c, ncols, r.reshape(-1, 1), v)
    return [m0, m1], mean, len(v)


def sparse_mode(idx, vals, mode, mode_size):
    """SparseMode::init: idx is nnz x nmodes; stable grouping by `mode`."""
    idx = np.asarray(idx, dtype=np.int64)
    others = np.delete(idx, mode, axis=1)
    return _group(idx[:, mode], mode_size, others, np.asarray(vals, dtype=np.float64))


def tensor_rows(idx, vals, dims):
    idx = np.asarray(idx, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    return [sparse_mode(idx, vals, m, dims[m]) for m in range(len(dims))], vals.mean(), len(vals)


def matrix_test(rows, cols, vals, nrows, ncols):
    """getTestData order for MatrixData: CSC of Ytest, duplicates summed."""
    r, c, v = sum_duplicates(rows, cols, vals, nrows, ncols)
    return np.stack([r, c], axis=1).astype(np.int32), v


def tensor_test(idx, vals, dims):
    """getTestData order for TensorData: SparseMode on mode 0."""
    idx = np.asarray(idx, dtype=np.int64)
    order = np.argsort(idx[:, 0], kind="stable")
    return np.ascontiguousarray(idx[order].astype(np.int32)), np.asarray(vals, dtype=np.float64)[order]


class Csr:
    def __init__(self, row_ptr, cols, vals, nrow, ncol):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.vals = None if vals is None else np.ascontiguousarray(vals, dtype=np.float64)
        self.nrow, self.ncol = nrow, ncol

    def c(self):
        return OrcCsr(self.nrow, self.ncol, _ptr(self.row_ptr), _ptr(self.cols), _ptr(self.vals))


def csr_from_coo(rows, cols, vals, nrow, ncol):
    """new_csr (libfastsparse): rows grouped, insertion order kept within a row."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    order = np.argsort(rows, kind="stable")
    row_ptr = np.zeros(nrow + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=nrow), out=row_ptr[1:])
    v = None if vals is None else np.asarray(vals, dtype=np.float64)[order]
    return Csr(row_ptr, cols[order], v, nrow, ncol)


def csr_pair(rows, cols, vals, nrow, ncol):
    """SparseDoubleFeat: F as CSR and its transpose as CSR (linop.h:39-42)."""
    return csr_from_coo(rows, cols, vals, nrow, ncol), csr_from_coo(cols, rows, vals, ncol, nrow)


# ---------------------------------------------------------------- numerics
def sample_latents(K, rows, mean, others, alpha, mu, Lambda, seed, it, mode, mu_rows=None):
    """Per-row sampler.  Factor matrices are K x N, passed as (N, K) C-order arrays."""
    N = rows.nrows
    U = np.zeros((N, K))
    oth = [np.ascontiguousarray(o, dtype=np.float64) for o in others]
    arr = (ctypes.c_void_p * len(oth))(*[_ptr(o) for o in oth])
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    Lam = np.asfortranarray(Lambda, dtype=np.float64)
    mr = None if mu_rows is None else np.ascontiguousarray(mu_rows, dtype=np.float64)
    rc = lib().orc_sample_latents(K, ctypes.byref(rows.c()), mean, ctypes.cast(arr, ctypes.c_void_p), alpha,
                                  _ptr(mu), _ptr(Lam), _ptr(mr), seed, it, mode, _ptr(U))
    if rc:
        raise RuntimeError("Cholesky Decomposition failed! (row %d)" % (-rc - 1))
    return U


def cond_normal_wishart(U, mu0, kappa, T, nu, seed, it, mode):
    """U as (N, K) C-order (i.e. K x N column-major)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    N, K = U.shape
    mu = np.zeros(K)
    Lam = np.zeros((K, K), order="F")
    rc = lib().orc_cond_normal_wishart(K, N, _ptr(U), _ptr(np.ascontiguousarray(mu0, dtype=np.float64)), kappa,
                                       _ptr(np.asfortranarray(T, dtype=np.float64)), nu, seed, it, mode,
                                       _ptr(mu), _ptr(Lam))
    if rc:
        raise RuntimeError("CondNormalWishart failed (%d)" % rc)
    return mu, np.array(Lam)


def train_sse(K, rows0, mean, factors):
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    arr = (ctypes.c_void_p * len(fs))(*[_ptr(f) for f in fs])
    return lib().orc_train_sse(K, len(fs), ctypes.byref(rows0.c()), mean, ctypes.cast(arr, ctypes.c_void_p))


def noise_alpha(sse, nnz, var_total, is_tensor, alpha_max, seed, it):
    return lib().orc_noise_alpha(sse, nnz, var_total, int(is_tensor), alpha_max, seed, it)


def eval_rmse(K, test_idx, test_vals, mean, factors, n, pred, pred_var):
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    arr = (ctypes.c_void_p * len(fs))(*[_ptr(f) for f in fs])
    ti = np.ascontiguousarray(test_idx, dtype=np.int32)
    tv = np.ascontiguousarray(test_vals, dtype=np.float64)
    r1, ra = _dbl(), _dbl()
    lib().orc_eval_rmse(K, len(fs), len(tv), _ptr(ti), _ptr(tv), mean, ctypes.cast(arr, ctypes.c_void_p), n,
                        _ptr(pred), _ptr(pred_var), ctypes.byref(r1), ctypes.byref(ra))
    return r1.value, ra.value


def norm_cdf(x):
    """Phi(x) (truncnorm.cpp:5-7)."""
    return lib().orc_norm_cdf(float(x))


def inv_norm_cdf(p):
    """Phi^{-1}(p) (the reference's inv_norm_cdf, truncnorm.cpp / tests.cpp:560-568)."""
    return lib().orc_inv_norm_cdf(float(p))


def truncnorm0(mean, seed, c0, it, tg):
    """One draw of rand_truncnorm(mean, 1, 0) (truncnorm.cpp:9-37) keyed by
    Philox block (c0, attempt, it, tg)."""
    return lib().orc_truncnorm0(float(mean), seed, c0, it, tg)


def probit_z(K, rows, others, U, seed, it, mode):
    """Augmented observations z (mode order) for probit sampling of one mode
    (latentprior.cpp:393-398); U is this mode's current factor (N, K)."""
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in others]
    arr = (ctypes.c_void_p * len(fs))(*[_ptr(f) for f in fs])
    U = np.ascontiguousarray(U, dtype=np.float64)
    z = np.zeros(int(rows.row_ptr[-1]) + 1)
    lib().orc_probit_z(K, ctypes.byref(rows.c()), ctypes.cast(arr, ctypes.c_void_p), _ptr(U), seed, it, mode,
                       _ptr(z))
    return z[:-1]


def auc(pred, test):
    """auc (bpmfutils.h:182-213)."""
    pred = np.ascontiguousarray(pred, dtype=np.float64)
    test = np.ascontiguousarray(test, dtype=np.float64)
    return lib().orc_auc(len(pred), _ptr(pred), _ptr(test))


def eval_probit(K, test_idx, test_vals, factors, n, pred, pred_var):
    """ProbitNoise::evalModel (noisemodels.cpp:129-159); returns (one-sample
    predictions, auc of them, auc of the running mean)."""
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    arr = (ctypes.c_void_p * len(fs))(*[_ptr(f) for f in fs])
    ti = np.ascontiguousarray(test_idx, dtype=np.int32)
    tv = np.ascontiguousarray(test_vals, dtype=np.float64)
    p1 = np.zeros(len(tv) + 1)
    a1, aa = _dbl(), _dbl()
    lib().orc_eval_probit(K, len(tv), _ptr(ti), _ptr(tv), ctypes.cast(arr, ctypes.c_void_p), n, _ptr(pred),
                          _ptr(pred_var), _ptr(p1), ctypes.byref(a1), ctypes.byref(aa))
    return p1[:-1], a1.value, aa.value


def sample_latents_uni(K, rows, mean, others, alpha, lam, mu_rows, U, seed, it, mode):
    """MacauOnePrior::sample_latents (macauoneprior.cpp:42-95, 108-187):
    coordinate-wise update of U (N x K, modified copy returned); lam is the
    diagonal prior precision, mu_rows (N x K) the per-row prior mean."""
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in others]
    arr = (ctypes.c_void_p * len(fs))(*[_ptr(f) for f in fs])
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    mr = np.ascontiguousarray(mu_rows, dtype=np.float64)
    lib().orc_sample_latents_uni(K, ctypes.byref(rows.c()), mean, ctypes.cast(arr, ctypes.c_void_p), alpha,
                                 _ptr(lam), _ptr(mr), seed, it, mode, _ptr(U))
    return U


def at_mul_bt_col(Ft, f, Z):
    """At_mul_Bt (linop.h:121-149): Z^T F[:, f] for Z (N x nr)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    y = np.zeros(Z.shape[1])
    lib().orc_at_mul_bt_col(Z.shape[1], ctypes.byref(Ft.c()), f, _ptr(Z), _ptr(y))
    return y


def add_acol_mul_bt(Ft, f, b, Z):
    """add_Acol_mul_bt (linop.h:151-178): returns Z + F[:, f] b^T."""
    Z = np.array(Z, dtype=np.float64, order="C", copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lib().orc_add_acol_mul_bt(Z.shape[1], ctypes.byref(Ft.c()), f, _ptr(b), _ptr(Z))
    return Z


def beta_sweep_uni(Ft, colsq, lam, lambda_beta, mu, U, Uhat, beta, seed, it, mode):
    """One feature sweep of MacauOnePrior::sample_beta (macauoneprior.cpp:215-252);
    returns the new beta (nfeat x K)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    N, K = U.shape
    beta = np.array(beta, dtype=np.float64, order="C", copy=True)
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (colsq, lam, lambda_beta, mu, Uhat)]
    lib().orc_beta_sweep_uni(K, N, ctypes.byref(Ft.c()), _ptr(args[0]), _ptr(args[1]), _ptr(args[2]), _ptr(args[3]),
                             _ptr(U), _ptr(args[4]), _ptr(beta), seed, it, mode)
    return beta


def lambda_beta_uni(beta, seed, it, mode):
    """sample_lambda_beta (macauoneprior.cpp:263-289)."""
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    out = np.zeros(beta.shape[1])
    lib().orc_lambda_beta_uni(beta.shape[1], beta.shape[0], _ptr(beta), seed, it, mode, _ptr(out))
    return out


def macau_one_update_prior(U, F, Ft, colsq, mu, lam, beta, Uhat, lambda_beta, seed, it, mode):
    """MacauOnePrior::update_prior (macauoneprior.cpp:189-289); updates and
    returns copies of (mu, lam, beta (nfeat x K), Uhat (N x K), lambda_beta)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    N, K = U.shape
    mu, lam = np.array(mu, dtype=np.float64), np.array(lam, dtype=np.float64)
    beta = np.array(beta, dtype=np.float64, order="C")
    Uhat = np.array(Uhat, dtype=np.float64, order="C")
    lb = np.array(lambda_beta, dtype=np.float64)
    cs = np.ascontiguousarray(colsq, dtype=np.float64)
    rc = lib().orc_macau_one_update_prior(K, mode, _ptr(U), N, ctypes.byref(F.c()), ctypes.byref(Ft.c()), _ptr(cs),
                                          _ptr(mu), _ptr(lam), _ptr(beta), _ptr(Uhat), _ptr(lb), seed, it)
    if rc:
        raise RuntimeError("MacauOnePrior update failed (%d)" % rc)
    return mu, lam, beta, Uhat, lb


def csr_A_mul_Bn(A, B):
    """out[r] = sum_c A[r,c] B[c]; B is (ncol, nrhs) C-order."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.zeros((A.nrow, B.shape[1]))
    lib().orc_csr_A_mul_Bn(B.shape[1], ctypes.byref(A.c()), _ptr(B), _ptr(out))
    return out


def AtA_mul_B(A, At, reg, B):
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.zeros_like(B)
    lib().orc_AtA_mul_B(B.shape[1], ctypes.byref(A.c()), ctypes.byref(At.c()), reg, _ptr(B), _ptr(out))
    return out


def blockcg(A, At, reg, B, tol, blocksize=None, excess=None):
    """Solve X (A^T A + reg I) = B; B is (nfeat, nrhs) C-order (= nrhs x nfeat col-major)."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    X = np.zeros_like(B)
    if blocksize is None:
        it = lib().orc_blockcg(B.shape[1], ctypes.byref(A.c()), ctypes.byref(At.c()), reg, _ptr(B), tol, _ptr(X))
    else:
        it = lib().orc_blockcg_blocked(B.shape[1], ctypes.byref(A.c()), ctypes.byref(At.c()), reg, _ptr(B), tol,
                                       blocksize, excess, _ptr(X))
    return X, it


def at_mul_a(A, At):
    """F^T F of a Csr (At_mul_A, linop.cpp:93-118): (ncol, ncol) with the lower
    triangle filled and the upper triangle zero."""
    out = np.zeros((A.ncol, A.ncol), order="F")
    lib().orc_at_mul_a(ctypes.byref(A.c()), ctypes.byref(At.c()), _ptr(out))
    return out


def ftf_solve(FtF, reg, B):
    """X (FtF + reg I) = B through Cholesky (latentprior.cpp:280-288); FtF is
    (n, n) with the lower triangle used, B is (n, nrhs) C-order (= nrhs x n
    col-major).  Raises when FtF + reg I is not positive definite."""
    FtF = np.asfortranarray(FtF, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    X = np.zeros_like(B)
    if lib().orc_ftf_solve(B.shape[1], B.shape[0], _ptr(FtF), reg, _ptr(B), _ptr(X)):
        raise RuntimeError("Cholesky decomp failed")
    return X


def sample_beta(U, mu, Lambda, F, Ft, lambda_beta, seed, it, mode=0, FtF=None, tol=1e-6):
    """MacauPrior::sample_beta (latentprior.cpp:272-293).  U is (N, K)
    C-order (K x N column-major); returns (beta as (nfeat, K) C-order, CG
    iteration count).  FtF (lower triangle, from at_mul_a) selects the F^T F
    branch."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    N, K = U.shape
    beta = np.zeros((F.ncol, K))
    it_out = ctypes.c_int(0)
    ftf = None if FtF is None else np.asfortranarray(FtF, dtype=np.float64)
    rc = lib().orc_sample_beta(K, mode, _ptr(U), N, _ptr(np.ascontiguousarray(mu, dtype=np.float64)),
                               _ptr(np.asfortranarray(Lambda, dtype=np.float64)), ctypes.byref(F.c()),
                               ctypes.byref(Ft.c()), _ptr(ftf), lambda_beta, tol, seed, it, _ptr(beta),
                               ctypes.byref(it_out))
    if rc:
        raise RuntimeError("sample_beta failed (%d)" % rc)
    return beta, it_out.value


def macau_update_prior(U, F, Ft, beta, Uhat, lambda_beta, seed, it, mode=0, FtF=None, tol=1e-6):
    """MacauPrior::update_prior (latentprior.cpp:254-264) with the prior
    constants of MacauPrior::init (WI = I, df = K; latentprior.cpp:189-195).
    U, Uhat are (N, K), beta (nfeat, K), all C-order.  Returns
    (mu, Lambda, beta, Uhat, lambda_beta, cg_iters) after the update."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    N, K = U.shape
    beta = np.array(beta, dtype=np.float64, order="C", copy=True)
    Uhat = np.array(Uhat, dtype=np.float64, order="C", copy=True)
    WI = np.eye(K, order="F")
    lb = _dbl(lambda_beta)
    mu = np.zeros(K)
    Lam = np.zeros((K, K), order="F")
    it_out = ctypes.c_int(0)
    ftf = None if FtF is None else np.asfortranarray(FtF, dtype=np.float64)
    rc = lib().orc_macau_update_prior(K, mode, _ptr(U), N, ctypes.byref(F.c()), ctypes.byref(Ft.c()), _ptr(ftf),
                                      _ptr(WI), K, _ptr(beta), _ptr(Uhat), ctypes.byref(lb), tol, _ptr(mu),
                                      _ptr(Lam), seed, it, ctypes.byref(it_out))
    if rc:
        raise RuntimeError("MacauPrior update failed (%d)" % rc)
    return mu, np.array(Lam), beta, Uhat, lb.value, it_out.value


def posterior_lambda_beta(beta, Lambda, nu, mu):
    """beta is (nfeat, K) C-order (K x nfeat col-major)."""
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    nfeat, K = beta.shape
    a, b = _dbl(), _dbl()
    lib().orc_posterior_lambda_beta(K, nfeat, _ptr(beta), _ptr(np.asfortranarray(Lambda, dtype=np.float64)),
                                    nu, mu, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


class RunResult:
    pass


def run(train_modes, mean, nnz, dims, K, test_idx, test_vals, *, is_tensor=False, adaptive=False,
        precision=1.0, sn_init=1.0, sn_max=10.0, side=None, lambda_beta=5.0, tol=1e-6, burnin=10,
        nsamples=10, seed=0, max_iters=0, want_U=False, use_ftf=None, probit=False, uni=None):
    """Full Gibbs run (MacauX::run).  want_U: also return the last
    iteration's U, mu and Lambda of every mode.  side: per-mode (F, Ft) Csr pairs or None;
    use_ftf: per-mode flags selecting the F^T F branch of sample_beta;
    probit: ProbitNoise (matrix data), reporting auc / auc1;
    uni: per-mode flags selecting MacauOnePrior for modes with side information."""
    M = len(dims)
    cfg = OrcRunCfg()
    cfg.nmodes, cfg.K, cfg.is_tensor = M, K, int(is_tensor)
    keep = []
    for m in range(M):
        cfg.dims[m] = dims[m]
        cfg.train[m] = train_modes[m].c()
    cfg.mean, cfg.nnz = mean, nnz
    ti = np.ascontiguousarray(test_idx, dtype=np.int32).reshape(-1, M)
    tv = np.ascontiguousarray(test_vals, dtype=np.float64)
    keep += [ti, tv]
    cfg.ntest, cfg.test_idx, cfg.test_vals = len(tv), _ptr(ti), _ptr(tv)
    cfg.noise_adaptive, cfg.precision, cfg.sn_init, cfg.sn_max = int(adaptive), precision, sn_init, sn_max
    cfg.noise_probit = int(bool(probit))
    side = side or [None] * M
    for m in range(M):
        if side[m] is not None:
            cfg.macau[m] = 1
            cfg.F[m] = side[m][0].c()
            cfg.Ft[m] = side[m][1].c()
            cfg.use_ftf[m] = int(bool(use_ftf and use_ftf[m]))
            cfg.uni[m] = int(bool(uni and uni[m]))
    cfg.lambda_beta0, cfg.tol = lambda_beta, tol
    cfg.burnin, cfg.nsamples, cfg.seed, cfg.max_iters = burnin, nsamples, seed, max_iters
    pred = np.zeros(len(tv))
    pvar = np.zeros(len(tv))
    total = burnin + nsamples if max_iters <= 0 else min(max_iters, burnin + nsamples)
    at = np.zeros(max(total, 1))
    rt = np.zeros(max(total, 1))
    lb = np.zeros(MAX_MODES)
    cfg.pred, cfg.pred_var, cfg.alpha_trace, cfg.rmse_trace = _ptr(pred), _ptr(pvar), _ptr(at), _ptr(rt)
    cfg.lambda_beta_out = _ptr(lb)
    Us, mus, Lams = [], [], []
    if want_U:
        for m in range(M):
            u = np.zeros((dims[m], K))
            Us.append(u)
            cfg.U_out[m] = _ptr(u)
            mus.append(np.zeros(K))
            Lams.append(np.zeros((K, K), order="F"))
            cfg.mu_out[m] = _ptr(mus[m])
            cfg.Lambda_out[m] = _ptr(Lams[m])
    rc = lib().orc_run(ctypes.byref(cfg))
    if rc:
        raise RuntimeError("oracle run failed (%d)" % rc)
    res = RunResult()
    res.rmse, res.rmse1, res.alpha = cfg.rmse, cfg.rmse1, cfg.alpha
    res.auc, res.auc1 = cfg.auc, cfg.auc1
    res.pred, res.pred_var = pred, pvar
    res.alpha_trace, res.rmse_trace = at[:total], rt[:total]
    res.U = Us
    res.mu, res.Lambda = mus, [np.array(L) for L in Lams]
    res.lambda_beta = lb[:M]
    res.cg_iters_last = cfg.cg_iters_last
    return res
