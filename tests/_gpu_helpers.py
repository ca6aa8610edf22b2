"""Helpers shared by the GPU parity tests: call the HIP path through the C ABI
and compare with the FP64 oracle on identical inputs and draws."""
import ctypes
import sys

import numpy as np

from macau_amd import _lib
from macau_amd._lib import check, ptr

import oracle as O

SEED = 20241202


def announce(msg):
    """Name the case before a GPU step (flushed, so it survives a crash)."""
    sys.stdout.write("[gpu-case] %s\n" % msg)
    sys.stdout.flush()


def f32(a):
    """Round to FP32 and back: the GPU stores factors in FP32, so the oracle
    gets the same input values."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def gpu_sample_latents(K, rows, mean, others, alpha, mu, Lam, seed, it, mode, sampler, mu_rows=None):
    L = _lib.lib()
    nrows = rows.nrows
    oth = [np.ascontiguousarray(o, dtype=np.float64) for o in others]
    arr = (ctypes.c_void_p * len(oth))(*[ptr(o) for o in oth])
    on = np.array([o.shape[0] for o in oth], dtype=np.int32)
    rp = np.ascontiguousarray(rows.row_ptr, dtype=np.int64)
    ix = np.ascontiguousarray(rows.idx, dtype=np.int32)
    vals = np.ascontiguousarray(rows.vals, dtype=np.float64)
    out = np.zeros((nrows, K))
    mr = None if mu_rows is None else np.ascontiguousarray(mu_rows, dtype=np.float64)
    check(L.macau_sample_latents(K, len(oth), nrows, ptr(rp), ptr(ix), ptr(vals), float(mean),
                                 ctypes.cast(arr, ctypes.c_void_p), ptr(on), float(alpha),
                                 ptr(np.ascontiguousarray(mu, dtype=np.float64)),
                                 ptr(np.asfortranarray(Lam, dtype=np.float64)), ptr(mr), seed, it, mode, sampler,
                                 ptr(out)))
    return out


def gpu_cond_normal_wishart(U, mu0, kappa, T, nu, seed, it, mode):
    U = np.ascontiguousarray(U, dtype=np.float64)
    N, K = U.shape
    mu = np.zeros(K)
    Lam = np.zeros((K, K), order="F")
    check(_lib.lib().macau_cond_normal_wishart(K, N, ptr(U), ptr(np.ascontiguousarray(mu0, dtype=np.float64)),
                                               kappa, ptr(np.asfortranarray(T, dtype=np.float64)), nu, seed, it,
                                               mode, ptr(mu), ptr(Lam)))
    return mu, np.array(Lam)


def gpu_blockcg(rows, cols, vals, nrow, ncol, reg, B, tol):
    """B is nrhs x nfeat; returns (X nrhs x nfeat, iterations)."""
    nrhs = B.shape[0]
    Bf = np.asfortranarray(B, dtype=np.float64)
    X = np.zeros((nrhs, ncol), order="F")
    it = ctypes.c_int(0)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.float64)
    check(_lib.lib().macau_solve_blockcg(nrhs, nrow, ncol, len(r), ptr(r), ptr(c), ptr(v), float(reg), ptr(Bf),
                                         float(tol), ptr(X), ctypes.byref(it)))
    return np.array(X), it.value


def row_rel_err(a, b):
    """max over rows of ||a_n - b_n|| / ||b_n||."""
    num = np.linalg.norm(a - b, axis=1)
    den = np.maximum(np.linalg.norm(b, axis=1), 1e-30)
    return float(np.max(num / den)) if len(num) else 0.0


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-30))


def spd(rng, K, scale=1.0):
    A = rng.standard_normal((K, K))
    return scale * (A @ A.T / K + np.eye(K))


def power_law_rows(rng, nrows, ncols, nnz_target, heavy=()):
    """Random rows with Zipf-like degrees; `heavy` lists explicit degrees of the first rows."""
    deg = np.minimum(ncols, np.maximum(0, (rng.pareto(1.2, nrows) * nnz_target / nrows).astype(int)))
    for i, d in enumerate(heavy):
        deg[i] = min(d, ncols)
    r = np.repeat(np.arange(nrows), deg)
    c = np.concatenate([rng.choice(ncols, d, replace=False) for d in deg]) if len(r) else np.zeros(0, int)
    y = rng.standard_normal(len(r)) + 3.0
    return r, c, y


class GpuModel:
    """Thin owner of a native model for the model-level parity tests."""

    def __init__(self, nmodes, K, precision, sn_max=10.0):
        L = _lib.lib()
        if precision == "adaptive":
            self.h = L.macau_make_adaptive(nmodes, K, 1.0, sn_max)
        elif precision == "probit":
            self.h = L.macau_make_probit(nmodes, K)
        else:
            self.h = L.macau_make_fixed(nmodes, K, float(precision))
        assert self.h, _lib.last_error()
        self.K, self.nmodes = K, nmodes

    def __del__(self):
        if getattr(self, "h", None):
            _lib.lib().macau_free(self.h)

    def data(self, idx, vals, dims, test_idx=None, test_vals=None):
        L = _lib.lib()
        idx = np.asfortranarray(idx, dtype=np.int32)
        d = np.array(dims, dtype=np.int32)
        check(L.macau_set_relation_data(self.h, ptr(idx), len(dims), ptr(np.ascontiguousarray(vals, dtype=np.float64)),
                                        idx.shape[0], ptr(d)))
        if test_idx is not None and len(test_vals):
            ti = np.asfortranarray(test_idx, dtype=np.int32)
            check(L.macau_set_relation_data_test(self.h, ptr(ti), len(dims),
                                                 ptr(np.ascontiguousarray(test_vals, dtype=np.float64)),
                                                 ti.shape[0], ptr(d)))

    def prior(self, side=None, lambda_beta=5.0, tol=1e-6, max_ff_size=None):
        """max_ff_size None: block CG always (macau_add_macau_prior); else the
        F^T F branch for side information with at most that many columns."""
        L = _lib.lib()
        if side is None:
            check(L.macau_add_bpmf_prior(self.h))
        else:
            r, c, v, nr, nc = side
            r = np.ascontiguousarray(r, dtype=np.int32)
            c = np.ascontiguousarray(c, dtype=np.int32)
            v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
            if max_ff_size is None:
                check(L.macau_add_macau_prior(self.h, nr, nc, len(r), ptr(r), ptr(c), ptr(v), lambda_beta, tol))
            else:
                check(L.macau_add_macau_prior_ff(self.h, nr, nc, len(r), ptr(r), ptr(c), ptr(v), lambda_beta, tol,
                                                 max_ff_size))

    def one_prior(self, side, lambda_beta=5.0):
        """MacauOnePrior (univariate) on COO side information (r, c, v or None, nrow, ncol)."""
        r, c, v, nr, nc = side
        r = np.ascontiguousarray(r, dtype=np.int32)
        c = np.ascontiguousarray(c, dtype=np.int32)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        check(_lib.lib().macau_add_macau_one_prior(self.h, nr, nc, len(r), ptr(r), ptr(c), ptr(v), lambda_beta))

    def setup(self, burnin, nsamples, seed, sampler, graph=True):
        L = _lib.lib()
        check(L.macau_set_samples(self.h, burnin, nsamples))
        check(L.macau_set_seed(self.h, seed))
        check(L.macau_set_sampler(self.h, sampler))
        check(L.macau_set_graph(self.h, int(graph)))
        check(L.macau_set_verbose(self.h, 0))
        check(L.macau_init(self.h))

    def iterate(self, n):
        L = _lib.lib()
        check(L.macau_iterate(self.h, n))
        check(L.macau_synchronize(self.h))

    def latents(self, mode, n):
        out = np.zeros((n, self.K))
        check(_lib.lib().macau_get_latents(self.h, mode, ptr(out)))
        return out

    def hyper(self, mode):
        mu = np.zeros(self.K)
        Lam = np.zeros((self.K, self.K), order="F")
        check(_lib.lib().macau_get_hyper(self.h, mode, ptr(mu), ptr(Lam)))
        return mu, np.array(Lam)

    def alpha(self):
        return _lib.lib().macau_get_noise_precision(self.h)

    def rmse(self):
        return _lib.lib().macau_get_rmse_avg(self.h)

    def auc(self):
        """(one-sample AUC, AUC of the averaged predictions) of a probit model."""
        a1, aa = ctypes.c_double(), ctypes.c_double()
        check(_lib.lib().macau_get_auc(self.h, ctypes.byref(a1), ctypes.byref(aa)))
        return a1.value, aa.value

    def link(self, mode, nfeat):
        b = np.zeros((nfeat, self.K))
        check(_lib.lib().macau_get_link(self.h, mode, ptr(b)))
        return b

    def predictions(self, n):
        p = np.zeros(n)
        check(_lib.lib().macau_get_predictions(self.h, ptr(p)))
        return p
