"""Python entry points of the sampler, mirroring jaak-s/macau's Cython module.

Same names, keyword arguments, defaults, errors and result layout as
``python/macau/macau.pyx`` (``macau`` :337-451, ``bpmf`` :221-234,
``Data`` :242-306, ``make_train_test`` :109-130, ``make_train_test_df``
:132-152, ``MacauResult`` :12-22, ``blockcg`` :67-91).  The sampling step
runs on the GPU through the C ABI of ``libmacau_amd.so``.

Differences from the reference, all additive:
  * ``seed=`` keyword: the reference seeds from the wall clock
    (macau.cpp:73); ``seed=None`` keeps that behaviour.
  * ``sampler=`` keyword: ``"auto"`` (default), ``"batched"`` (split-BF16
    MFMA Gram, nnz-bucketed) or ``"plain"`` (FP32, one work-group per row).
Side information with at most 10000 features solves the link matrix through
the dense F'F Cholesky branch (latentprior.cpp:280-289), larger side
information through the block CG, as in the reference.  ``precision="probit"``
runs ProbitNoise on matrix data; ``univariate=True`` puts MacauOnePrior on
the modes with side information, as the reference does.
"""
import ctypes
import numbers
import signal
import sys
import time

import numpy as np
import pandas as pd
import scipy.sparse

from . import _lib
from ._lib import check, ptr

_SAMPLERS = {"plain": _lib.SAMPLER_PLAIN_FP32, "batched": _lib.SAMPLER_BATCHED}


def _pick_sampler(sampler, num_latent, nmodes):
    """'auto' takes the batched kernels where they exist for this latent size."""
    if sampler == "auto":
        ok = _lib.lib().macau_sampler_supported(_lib.SAMPLER_BATCHED, num_latent, nmodes)
        return "batched" if ok else "plain"
    return sampler


class MacauResult(object):
    def __init__(self):
        pass

    def __repr__(self):
        s = ("Matrix factorization results\n" +
             "Test RMSE:        %.4f\n" % self.rmse_test +
             "Matrix size:      [%s]\n" % " x ".join([str(x) for x in self.Yshape]) +
             "Number of train:  %d\n" % self.ntrain +
             "Number of test:   %d\n" % self.ntest +
             "To see predictions on test set see '.prediction' field.")
        return s


def _is_sparse_matrix(Y):
    return scipy.sparse.issparse(Y) and Y.format in ("coo", "csr", "csc")


def make_train_test(Y, ntest):
    """Random train/test partition of the observed cells of a sparse matrix
    (macau.pyx:109-130).

    Y is a scipy COO, CSR or CSC matrix.  ntest below 1 is the fraction of
    Y's nonzeros that go to the test side, ntest of 1 or more their count
    (rounded).  The cells are drawn with numpy's global generator; both
    results are COO matrices of Y's shape, train first.
    """
    if not _is_sparse_matrix(Y):
        raise TypeError("Unsupported Y type: %s" % type(Y))
    if not isinstance(ntest, numbers.Real) or ntest < 0:
        raise TypeError("ntest has to be a non-negative number (number or ratio of test samples).")
    Y = Y.tocoo(copy=False)
    if ntest < 1:
        ntest = Y.nnz * ntest
    ntest = int(round(ntest))
    rperm = np.random.permutation(Y.nnz)
    train = rperm[ntest:]
    test = rperm[0:ntest]
    Ytrain = scipy.sparse.coo_matrix((Y.data[train], (Y.row[train], Y.col[train])), shape=Y.shape)
    Ytest = scipy.sparse.coo_matrix((Y.data[test], (Y.row[test], Y.col[test])), shape=Y.shape)
    return Ytrain, Ytest


def make_train_test_df(Y, ntest):
    """Random train/test partition of the rows of a DataFrame in the cell
    layout (index columns, then the value column; macau.pyx:132-152).

    ntest below 1 is the fraction of rows that go to the test side, ntest of
    1 or more their count (rounded).  Returns (train, test) as row subsets of
    Y, in the permuted order of numpy's global generator.
    """
    if type(Y) != pd.core.frame.DataFrame:
        raise TypeError("Y should be DataFrame.")
    if not isinstance(ntest, numbers.Real) or ntest < 0:
        raise TypeError("ntest has to be a non-negative number (number or ratio of test samples).")
    if ntest < 1:
        ntest = Y.shape[0] * ntest
    ntest = int(round(ntest))
    rperm = np.random.permutation(Y.shape[0])
    train = rperm[ntest:]
    test = rperm[0:ntest]
    return Y.iloc[train], Y.iloc[test]


def remove_nan(Y):
    if not np.any(np.isnan(Y.data)):
        return Y
    idx = np.where(np.isnan(Y.data) == False)[0]  # noqa: E712
    return scipy.sparse.coo_matrix((Y.data[idx], (Y.row[idx], Y.col[idx])), shape=Y.shape)


class Data:
    """Training/test data in the reference's layout (macau.pyx:242-306)."""

    def __init__(self, Y, Ytest):
        if _is_sparse_matrix(Y):
            if Ytest is None:
                Ytest = scipy.sparse.coo_matrix(Y.shape, dtype=np.float64)
            if isinstance(Ytest, numbers.Real):
                Y, Ytest = make_train_test(Y, Ytest)
            if not _is_sparse_matrix(Ytest):
                raise ValueError("When Y is a sparse matrix Ytest must be too.")
            if Y.shape != Ytest.shape:
                raise ValueError("Y (%d x %d) and Ytest (%d x %d) must have the same shape." %
                                 (Y.shape[0], Y.shape[1], Ytest.shape[0], Ytest.shape[1]))
            Y = remove_nan(Y.tocoo(copy=False))
            Ytest = remove_nan(Ytest.tocoo(copy=False))
            self.shape = np.array(Y.shape, dtype=np.int32)
            self.idxTrain = [Y.row, Y.col]
            self.valTrain = np.asarray(Y.data, dtype=np.float64)
            self.idxTest = [Ytest.row, Ytest.col]
            self.valTest = np.asarray(Ytest.data, dtype=np.float64)
            self.colnames = np.array(["row", "col"], dtype=object)
        elif type(Y) == pd.core.frame.DataFrame:
            if Ytest is None:
                Ytrain = Y
                Ytest = Y[0:0]
            if isinstance(Ytest, numbers.Real):
                Ytrain, Ytest = make_train_test_df(Y, Ytest)
            else:
                Ytrain = Y
            if type(Ytest) != pd.core.frame.DataFrame:
                raise TypeError("When Y is a DataFrame Ytest must be too.")
            if (Y.columns != Ytest.columns).any():
                raise ValueError("Columns of Y and Ytest must be the same.")
            if (Y.dtypes != Ytest.dtypes).any():
                raise TypeError("Y.dtypes and Ytest.dtypes must be the same.")
            int_cols = [c for c in Ytrain.columns if Ytrain[c].dtype in (np.int64, np.int32)]
            float_cols = [c for c in Ytrain.columns if Ytrain[c].dtype in (np.float32, np.float64)]
            if len(int_cols) > 10:
                raise ValueError("Y has too many index(int) columns (%d), maximum is 10." % len(int_cols))
            if len(int_cols) < 2:
                raise ValueError("Y must have at least 2 index (int) columns.")
            if len(float_cols) != 1:
                raise ValueError("Y has %d float columns but must have exactly 1 value column." % len(float_cols))
            if Ytrain.shape[0] == 0:
                raise ValueError("Ytrain must not be empty.")
            value_col = float_cols[0]
            self.colnames = int_cols
            if Ytest.shape[0] > 0:
                self.shape = np.array(np.maximum([Y[c].max() for c in int_cols],
                                                 [Ytest[c].max() for c in int_cols]) + 1, dtype=np.int32)
            else:
                self.shape = np.array([Y[c].max() for c in int_cols], dtype=np.int32) + 1
            # the reference indexes with Y (not Ytrain) here (macau.pyx:300-303)
            self.idxTrain = [np.array(Y[c], dtype=np.int32) for c in int_cols]
            self.idxTest = [np.array(Ytest[c], dtype=np.int32) for c in int_cols]
            self.valTrain = np.array(Y[value_col], dtype=np.float64)
            self.valTest = np.array(Ytest[value_col], dtype=np.float64)
        else:
            raise TypeError("Unsupported Y type: %s" % type(Y))


def _idx_matrix(idx_list):
    idx = np.zeros([len(idx_list[0]), len(idx_list)], dtype=np.int32, order="F")
    for i in range(len(idx_list)):
        idx[:, i] = idx_list[i]
    return idx


class Model:
    """Owner of one native model handle (macau.h:11-54)."""

    def __init__(self, nmodes, num_latent, precision, sn_max=10.0):
        L = _lib.lib()
        if isinstance(precision, str):
            if precision == "adaptive" or precision == "sample":
                h = L.macau_make_adaptive(nmodes, num_latent, 1.0, float(sn_max))
            elif precision == "probit":
                h = L.macau_make_probit(nmodes, num_latent)
            else:
                raise ValueError("Parameter 'precision' has to be either a number or \"adaptive\" for adaptive "
                                 "precision, or \"probit\" for binary matrices.")
        else:
            h = L.macau_make_fixed(nmodes, num_latent, float(precision))
        if not h:
            raise RuntimeError(_lib.last_error())
        self.h = h
        self.nmodes = nmodes
        self.num_latent = num_latent

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            _lib.lib().macau_free(h)
            self.h = None

    def set_data(self, data):
        dims = np.array(data.shape, dtype=np.int32)
        idx = _idx_matrix(data.idxTrain)
        vals = np.ascontiguousarray(data.valTrain, dtype=np.float64)
        check(_lib.lib().macau_set_relation_data(self.h, ptr(idx), len(dims), ptr(vals), idx.shape[0], ptr(dims)))
        te_idx = _idx_matrix(data.idxTest)
        te_vals = np.ascontiguousarray(data.valTest, dtype=np.float64)
        if te_idx.shape[0] > 0:
            check(_lib.lib().macau_set_relation_data_test(self.h, ptr(te_idx), len(dims), ptr(te_vals),
                                                          te_idx.shape[0], ptr(dims)))

    def add_prior(self, side, lambda_beta, tol, max_ff_size=10000):
        """make_prior (macau.pyx:154-192); side information with at most
        max_ff_size columns solves the link matrix through F^T F (compute_ff,
        macau.pyx:160), like setSidePriors (macau.pyx:334)."""
        L = _lib.lib()
        if side is None:
            check(L.macau_add_bpmf_prior(self.h))
            return
        if not (_is_sparse_matrix(side) or type(side) == np.ndarray):
            raise TypeError("Unsupported side information type: '%s'" % type(side).__name__)
        if type(side) == np.ndarray:
            if len(side.shape) != 2:
                raise TypeError("Side information must have 2 dimensions (got %d)." % len(side.shape))
            X = np.asarray(side, dtype=np.float64)
            r, c = np.nonzero(np.ones(X.shape, dtype=bool))
            v = X[r, c]
        else:
            X = side.tocoo(copy=False)
            r, c, v = X.row, X.col, X.data
        r = np.ascontiguousarray(r, dtype=np.int32)
        c = np.ascontiguousarray(c, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(L.macau_add_macau_prior_ff(self.h, side.shape[0], side.shape[1], len(v), ptr(r), ptr(c), ptr(v),
                                         float(lambda_beta), float(tol), int(max_ff_size)))

    def add_one_prior(self, side, lambda_beta):
        """make_one_prior (macau.pyx:194-214): BPMF for no side information,
        else the univariate MacauOnePrior on sparse side information."""
        L = _lib.lib()
        if (side is None) or (isinstance(side, tuple) and side == ()):
            check(L.macau_add_bpmf_prior(self.h))
            return
        if not _is_sparse_matrix(side):
            raise ValueError("Unsupported side information type: '%s'" % type(side).__name__)
        X = side.tocoo(copy=False)
        r = np.ascontiguousarray(X.row, dtype=np.int32)
        c = np.ascontiguousarray(X.col, dtype=np.int32)
        v = np.ascontiguousarray(X.data, dtype=np.float64)
        binary = bool(np.all(v == 1))
        check(L.macau_add_macau_one_prior(self.h, side.shape[0], side.shape[1], len(v), ptr(r), ptr(c),
                                          None if binary else ptr(v), float(lambda_beta)))

    def get_dims(self):
        d = np.zeros(self.nmodes, dtype=np.int32)
        check(_lib.lib().macau_get_dims(self.h, ptr(d)))
        return d

    def latents(self, mode):
        n = int(self.get_dims()[mode])
        out = np.zeros((n, self.num_latent))
        check(_lib.lib().macau_get_latents(self.h, mode, ptr(out)))
        return out

    def hyper(self, mode):
        K = self.num_latent
        mu = np.zeros(K)
        Lam = np.zeros((K, K), order="F")
        check(_lib.lib().macau_get_hyper(self.h, mode, ptr(mu), ptr(Lam)))
        return mu, np.array(Lam)


def bpmf(Y, Ytest=None, num_latent=10, precision=1.0, burnin=50, nsamples=400, **keywords):
    return macau(Y, Ytest=Ytest, num_latent=num_latent, precision=precision, burnin=burnin, nsamples=nsamples,
                 **keywords)


def macau(Y,
          Ytest=None,
          side=[],
          lambda_beta=5.0,
          num_latent=10,
          precision=1.0,
          burnin=50,
          nsamples=400,
          univariate=False,
          tol=1e-6,
          sn_max=10.0,
          save_prefix=None,
          verbose=True,
          seed=None,
          sampler="auto"):
    """Gibbs sampling of a matrix or tensor factorization, with side
    information on any mode (macau.pyx:337-451); returns a MacauResult.

    Y           observed training cells: a scipy sparse matrix, or a
                DataFrame whose last column holds the values and whose other
                columns (2 to 10) the integer index of each mode.
    Ytest       the test cells in the same form as Y, or a number below 1:
                that fraction of Y is held out at random as the test set.
    side        one entry per mode (rows first): None, a dense array or a
                scipy sparse matrix with as many rows as that mode has items.
    lambda_beta starting precision of the link matrix beta between a side
                matrix and the latents (sampled afterwards).
    num_latent  latent dimension K.
    precision   observation noise: a number (fixed precision), "adaptive"
                (precision sampled from the residuals, capped by sn_max) or
                "probit" (binary cells, latent Gaussian augmentation).
    burnin, nsamples
                iterations discarded, then averaged into the predictions.
    univariate  coordinate-wise (MacauOnePrior) updates for modes with side
                information instead of the joint beta draw.
    tol         relative residual at which the block CG solve of beta stops.
    sn_max      largest signal-to-noise ratio "adaptive" may reach.
    save_prefix write the posterior-mean latents, link matrices and priors
                under this prefix at the end (None: write nothing).
    verbose     print one status line per iteration.
    seed        key of the counter-based random streams (None: from the clock).
    sampler     "auto" (the split-BF16 MFMA kernels where they support
                num_latent, else the FP32 ones), "batched" or "plain".
    """
    data = Data(Y, Ytest)

    if not side:
        side = [None for i in range(len(data.shape))]
    if type(side) not in [list, tuple]:
        raise ValueError("Parameter 'side' must be a tuple or a list.")
    if len(side) != len(data.shape):
        raise ValueError("Length of 'side' is %d but must be equal to the number of data dimensions (%d)." %
                         (len(side), len(data.shape)))
    if sampler != "auto" and sampler not in _SAMPLERS:
        raise ValueError("sampler must be one of %s" % sorted(list(_SAMPLERS) + ["auto"]))

    D = int(np.int32(num_latent))
    nmodes = len(data.shape)
    if univariate and isinstance(precision, str) and precision == "probit":
        raise ValueError("Univariate sampler for probit model is not yet implemented.")
    model = Model(nmodes, D, precision, sn_max)
    L = _lib.lib()
    model.set_data(data)
    for s in side:
        if univariate:
            model.add_one_prior(s, np.float64(lambda_beta))
        else:
            model.add_prior(s, np.float64(lambda_beta), np.float64(tol))

    check(L.macau_set_samples(model.h, int(burnin), int(nsamples)))
    check(L.macau_set_verbose(model.h, int(bool(verbose))))
    if seed is None:
        seed = int(time.time() * 1e6) & 0xFFFFFFFFFFFFFFFF
    check(L.macau_set_seed(model.h, int(seed)))
    check(L.macau_set_sampler(model.h, _SAMPLERS[_pick_sampler(sampler, D, nmodes)]))

    if save_prefix is None:
        check(L.macau_set_save_model(model.h, 0))
    else:
        if type(save_prefix) != str:
            raise ValueError("Parameter 'save_prefix' has to be a string (str) or None.")
        check(L.macau_set_save_model(model.h, 1))
        check(L.macau_set_save_prefix(model.h, save_prefix.encode()))

    rc = L.macau_run(model.h)
    # restoring Python default signal handler
    signal.signal(signal.SIGINT, signal.default_int_handler)
    sys.stdout.flush()
    check(rc)

    ntest = int(L.macau_get_test_nonzeros(model.h))
    yhat = np.zeros(ntest)
    yhat_sd = np.zeros(ntest)
    testdata = np.zeros((ntest, nmodes + 1), order="F")
    check(L.macau_get_predictions(model.h, ptr(yhat)))
    check(L.macau_get_stds(model.h, ptr(yhat_sd)))
    check(L.macau_get_test_data(model.h, ptr(testdata)))

    df = pd.DataFrame(testdata[:, :-1], columns=data.colnames, dtype="int")
    df["y"] = pd.Series(testdata[:, -1])
    df["y_pred"] = pd.Series(yhat)
    df["y_pred_std"] = pd.Series(yhat_sd)

    result = MacauResult()
    if isinstance(precision, str) and precision == "probit":
        result.rmse_test = np.nan
        result.auc_test = L.macau_get_rmse_test(model.h)
    else:
        result.rmse_test = L.macau_get_rmse_test(model.h)
        result.auc_test = np.nan
    result.Yshape = data.shape
    result.ntrain = data.valTrain.shape[0]
    result.ntest = data.valTest.shape[0] if Ytest is not None else 0
    result.prediction = pd.DataFrame(df)
    del model
    return result


def blockcg(X, B, reg, tol=1e-6):
    """Block CG solve of B = beta (X'X + reg I) with a binary X (macau.pyx:67-91).

    B must be Fortran-ordered (nrhs x nfeat).  Returns [iterations, seconds]
    like the reference.
    """
    if not np.isfortran(B):
        raise ValueError("B must have order='F' (fortran)")
    if B.shape[1] != X.shape[1]:
        raise ValueError("B.shape[1] must equal X.shape[1]")
    X = X.tocoo(copy=False)
    rows = np.ascontiguousarray(X.row, dtype=np.int32)
    cols = np.ascontiguousarray(X.col, dtype=np.int32)
    result = np.zeros((B.shape[0], B.shape[1]), order="F")
    Bc = np.asfortranarray(B, dtype=np.float64)
    print("Sparse [%d x %d] matrix" % (X.shape[0], X.shape[1]))
    print("Running block-cg")
    iters = ctypes.c_int(0)
    start = time.perf_counter()
    check(_lib.lib().macau_solve_blockcg(B.shape[0], X.shape[0], X.shape[1], len(rows), ptr(rows), ptr(cols),
                                         None, float(reg), ptr(Bc), float(tol), ptr(result), ctypes.byref(iters)))
    end = time.perf_counter()
    v = np.zeros(2)
    v[0] = iters.value
    v[1] = end - start
    return v
