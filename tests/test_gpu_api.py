"""The Python entry points on the GPU, mirroring the reference's own tests
(python/macau/test_macau.py).  Every case runs the full Gibbs chain through
the HIP library; quality bounds are the reference's (tensor RMSE < 0.5 at
precision 50, test_macau.py:132-207; probit AUC > 0.55, test_macau.py:47-62;
univariate tensor RMSE < 0.5, test_macau.py:191-207)."""
import itertools
import os

import numpy as np
import pandas as pd
import pytest
import scipy.sparse

import macau_amd
from _gpu_helpers import announce

pytestmark = pytest.mark.gpu


def _rand(m, n, d, seed):
    return scipy.sparse.random(m, n, density=d, random_state=seed, format="coo")


def test_bpmf():  # test_macau.py:9-16
    announce("api bpmf")
    Y = _rand(10, 20, 0.2, 1)
    Y, Ytest = macau_amd.make_train_test(Y, 0.5)
    res = macau_amd.bpmf(Y, Ytest=Ytest, num_latent=4, verbose=False, burnin=50, nsamples=50,
                         univariate=False, seed=1)
    assert Ytest.nnz == res.prediction.shape[0]
    assert list(res.prediction.columns[0:2]) == ["row", "col"]
    assert np.all(np.isfinite(res.prediction.y_pred))
    assert np.all(res.prediction.y_pred_std >= 0)


def test_bpmf_numerictest():  # test_macau.py:18-21
    announce("api bpmf numeric test fraction")
    X = _rand(15, 10, 0.2, 2)
    res = macau_amd.bpmf(X, 0.3, num_latent=10, burnin=10, nsamples=15, verbose=False, seed=2)
    assert res.ntest == res.prediction.shape[0] > 0


def test_macau():  # test_macau.py:23-32
    announce("api macau dense-in-coo side on both modes")
    rng = np.random.default_rng(3)
    Y = _rand(10, 20, 0.2, 3)
    Y, Ytest = macau_amd.make_train_test(Y, 0.5)
    side1 = scipy.sparse.coo_matrix(rng.random((10, 2)))
    side2 = scipy.sparse.coo_matrix(rng.random((20, 3)))
    res = macau_amd.macau(Y, Ytest=Ytest, side=[side1, side2], num_latent=4, verbose=False, burnin=50,
                          nsamples=50, univariate=False, seed=3)
    assert Ytest.nnz == res.prediction.shape[0]
    assert np.isfinite(res.rmse_test)


def test_macau_side_bin():  # test_macau.py:34-39
    announce("api macau binary side")
    X = _rand(15, 10, 0.2, 4)
    Xt = _rand(15, 10, 0.1, 5)
    F = _rand(15, 2, 0.5, 6)
    F.data[:] = 1
    res = macau_amd.macau(X, Xt, side=[F, None], num_latent=5, burnin=10, nsamples=5, verbose=False, seed=4)
    assert res.prediction.shape[0] == Xt.nnz


def test_macau_dense():  # test_macau.py:41-45
    announce("api macau dense side")
    rng = np.random.default_rng(7)
    Y = _rand(15, 10, 0.2, 7)
    Yt = _rand(15, 10, 0.1, 8)
    F = rng.standard_normal((15, 2))
    res = macau_amd.macau(Y, Yt, side=[F, None], num_latent=5, burnin=10, nsamples=5, verbose=False, seed=5)
    assert np.isfinite(res.rmse_test)


def test_macau_dense_probit():  # test_macau.py:47-62
    announce("api macau dense probit")
    rng = np.random.default_rng(21)
    A = rng.standard_normal((25, 2))
    B = rng.standard_normal((3, 2))
    idx = list(itertools.product(np.arange(A.shape[0]), np.arange(B.shape[0])))
    df = pd.DataFrame(np.asarray(idx), columns=["A", "B"])
    df["value"] = (np.array([np.sum(A[i[0], :] * B[i[1], :]) for i in idx]) > 0.0).astype(np.float64)
    np.random.seed(21)
    Ytrain, Ytest = macau_amd.make_train_test_df(df, 0.2)
    results = macau_amd.macau(Y=Ytrain, Ytest=Ytest, side=[A, None], num_latent=4, verbose=False, burnin=20,
                              nsamples=20, univariate=False, precision="probit", seed=21)
    assert (results.prediction.columns[0:2] == ["A", "B"]).all()
    assert results.auc_test > 0.55, "AUC %f" % results.auc_test
    assert np.isnan(results.rmse_test)
    p = results.prediction["y_pred"].values
    assert np.all((p > 0) & (p < 1))


def test_bpmf_emptytest_probit():  # test_macau.py:84-88
    announce("api bpmf probit empty test set")
    X = _rand(15, 10, 0.2, 22)
    X.data = (X.data > 0.5).astype(np.float64)
    r0 = macau_amd.bpmf(X, Ytest=0, num_latent=10, burnin=10, nsamples=15, precision="probit", verbose=False,
                        seed=22)
    r1 = macau_amd.bpmf(X, Ytest=None, num_latent=10, burnin=10, nsamples=15, precision="probit", verbose=False,
                        seed=22)
    assert np.isnan(r0.auc_test) and np.isnan(r1.auc_test)
    assert r0.prediction.shape[0] == 0


def test_macau_univariate():  # test_macau.py:64-73
    announce("api univariate sampler")
    Y = _rand(10, 20, 0.2, 9)
    np.random.seed(9)
    Y, Ytest = macau_amd.make_train_test(Y, 0.5)
    side1 = scipy.sparse.coo_matrix(np.random.rand(10, 2))
    side2 = scipy.sparse.coo_matrix(np.random.rand(20, 3))
    results = macau_amd.bpmf(Y, Ytest=Ytest, side=[side1, side2], num_latent=4, verbose=False, burnin=50,
                             nsamples=50, univariate=True, seed=9)
    assert Ytest.nnz == results.prediction.shape[0]
    assert np.isfinite(results.rmse_test)


def test_macau_tensor_univariate():  # test_macau.py:191-207
    announce("api univariate tensor")
    rng = np.random.default_rng(23)
    A = rng.standard_normal((30, 2))
    B = rng.standard_normal((4, 2))
    C = rng.standard_normal((2, 2))
    idx = list(itertools.product(np.arange(A.shape[0]), np.arange(B.shape[0]), np.arange(C.shape[0])))
    df = pd.DataFrame(np.asarray(idx), columns=["A", "B", "C"])
    df["value"] = np.array([np.sum(A[i[0], :] * B[i[1], :] * C[i[2], :]) for i in idx])
    np.random.seed(23)
    Ytrain, Ytest = macau_amd.make_train_test_df(df, 0.2)
    Acoo = scipy.sparse.coo_matrix(A)
    results = macau_amd.macau(Y=Ytrain, Ytest=Ytest, side=[Acoo, None, None], num_latent=4, verbose=False,
                              burnin=20, nsamples=20, univariate=True, precision=50, seed=23)
    assert results.rmse_test < 0.5, "RMSE %f" % results.rmse_test


def test_univariate_errors():  # macau.pyx:194-199, 393-395
    Y = _rand(10, 20, 0.2, 9)
    with pytest.raises(ValueError):
        macau_amd.bpmf(Y, verbose=False, univariate=True, precision="probit", burnin=2, nsamples=2)
    with pytest.raises(ValueError):
        macau_amd.macau(Y, side=[np.ones((10, 2)), None], verbose=False, univariate=True, burnin=2, nsamples=2)


def test_bpmf_emptytest():  # test_macau.py:80-82
    announce("api bpmf empty test set")
    X = _rand(15, 10, 0.2, 10)
    r0 = macau_amd.bpmf(X, Ytest=0, num_latent=10, burnin=10, nsamples=15, verbose=False, seed=6)
    r1 = macau_amd.bpmf(X, Ytest=None, num_latent=10, burnin=10, nsamples=15, verbose=False, seed=6)
    assert np.isnan(r0.rmse_test) and np.isnan(r1.rmse_test)
    assert r0.prediction.shape[0] == 0


def test_bpmf_tensor():  # test_macau.py:114-130
    announce("api bpmf tiny tensor")
    rng = np.random.default_rng(1234)
    Y = pd.DataFrame({"A": rng.integers(0, 5, 7), "B": rng.integers(0, 4, 7), "C": rng.integers(0, 3, 7),
                      "value": rng.standard_normal(7)})
    Ytest = pd.DataFrame({"A": rng.integers(0, 5, 5), "B": rng.integers(0, 4, 5), "C": rng.integers(0, 3, 5),
                          "value": rng.standard_normal(5)})
    res = macau_amd.bpmf(Y, Ytest=Ytest, num_latent=4, verbose=False, burnin=50, nsamples=50, univariate=False,
                         seed=7)
    assert res.prediction.shape[0] == 5
    assert list(res.prediction.columns[0:3]) == ["A", "B", "C"]


def _cp_frame(dims, seed):
    rng = np.random.default_rng(seed)
    F = [rng.standard_normal((d, 2)) for d in dims]
    idx = np.asarray(list(itertools.product(*[np.arange(d) for d in dims])))
    cols = ["A", "B", "C"][:len(dims)]
    df = pd.DataFrame(idx, columns=cols)
    prod = np.ones((len(idx), 2))
    for m, f in enumerate(F):
        prod *= f[idx[:, m]]
    df["value"] = prod.sum(1)
    return df, F


def test_bpmf_tensor2():  # test_macau.py:132-146
    announce("api bpmf tensor2 quality")
    df, _ = _cp_frame((15, 20, 3), 11)
    Ytrain, Ytest = macau_amd.make_train_test_df(df, 0.2)
    res = macau_amd.bpmf(Y=Ytrain, Ytest=Ytest, num_latent=4, verbose=False, burnin=20, nsamples=20,
                         univariate=False, precision=50, seed=8)
    assert res.rmse_test < 0.5


def test_bpmf_tensor3():  # test_macau.py:148-168
    announce("api bpmf tensor3 quality and matrix form")
    df, _ = _cp_frame((15, 20, 1), 12)
    Ytrain, Ytest = macau_amd.make_train_test_df(df, 0.2)
    res = macau_amd.bpmf(Y=Ytrain, Ytest=Ytest, num_latent=4, verbose=False, burnin=20, nsamples=20,
                         univariate=False, precision=50, seed=9)
    assert res.rmse_test < 0.5
    Ytrain_sp = scipy.sparse.coo_matrix((Ytrain.value, (Ytrain.A, Ytrain.B)))
    Ytest_sp = scipy.sparse.coo_matrix((Ytest.value, (Ytest.A, Ytest.B)))
    res_mat = macau_amd.bpmf(Y=Ytrain_sp, Ytest=Ytest_sp, num_latent=4, verbose=False, burnin=20, nsamples=20,
                             univariate=False, precision=50, seed=9)
    assert np.isfinite(res_mat.rmse_test)


def test_macau_tensor():  # test_macau.py:170-189
    announce("api macau tensor with side on mode 0")
    df, F = _cp_frame((30, 4, 2), 13)
    Ytrain, Ytest = macau_amd.make_train_test_df(df, 0.2)
    Acoo = scipy.sparse.coo_matrix(F[0])
    res = macau_amd.macau(Y=Ytrain, Ytest=Ytest, side=[Acoo, None, None], num_latent=4, verbose=False, burnin=20,
                          nsamples=20, univariate=False, precision=50, seed=10)
    assert list(res.prediction.columns[0:3]) == ["A", "B", "C"]
    assert res.rmse_test < 0.5


def test_macau_tensor_empty():  # test_macau.py:208-222
    announce("api macau tensor empty test set")
    df, _ = _cp_frame((30, 4, 2), 14)
    r0 = macau_amd.macau(df, Ytest=0, num_latent=2, burnin=5, nsamples=5, precision=1.0, verbose=False, seed=11)
    r1 = macau_amd.macau(df, Ytest=None, num_latent=2, burnin=5, nsamples=5, precision=1.0, verbose=False, seed=11)
    assert np.isnan(r0.rmse_test) and np.isnan(r1.rmse_test)


@pytest.mark.parametrize("sampler", ["plain", "batched", "auto"])
def test_samplers_agree(sampler):
    announce("api sampler=%s" % sampler)
    rng = np.random.default_rng(15)
    U, V = rng.standard_normal((300, 3)), rng.standard_normal((200, 3))
    Y = _rand(300, 200, 0.1, 15)
    Y.data = (U[Y.row] * V[Y.col]).sum(1) + 0.1 * rng.standard_normal(Y.nnz)
    np.random.seed(15)  # make_train_test draws its split from numpy's global stream
    Ytr, Yte = macau_amd.make_train_test(Y, 0.2)
    res = macau_amd.bpmf(Ytr, Ytest=Yte, num_latent=32, burnin=20, nsamples=20, precision=100, verbose=False,
                         seed=16, sampler=sampler)
    plain = macau_amd.bpmf(Ytr, Ytest=Yte, num_latent=32, burnin=20, nsamples=20, precision=100, verbose=False,
                           seed=16, sampler="plain")
    assert abs(res.rmse_test / plain.rmse_test - 1.0) < 5e-3
    assert res.rmse_test < 0.3


def test_save_model(tmp_path):
    announce("api save_prefix model files")
    prefix = str(tmp_path / "m")
    X = _rand(20, 15, 0.3, 17)
    F = _rand(20, 6, 0.5, 18)
    macau_amd.macau(X, 0.2, side=[F, None], num_latent=3, burnin=2, nsamples=2, verbose=False, save_prefix=prefix,
                    seed=12)
    names = sorted(os.listdir(tmp_path))
    assert "m-meanvalue.csv" in names
    for s in (1, 2):
        for u in (1, 2):
            assert "m-sample%d-U%d-latents.csv" % (s, u) in names
            assert "m-sample%d-U%d-latentmean.csv" % (s, u) in names
        assert "m-sample%d-U1-link.csv" % s in names
    lat = np.loadtxt(tmp_path / "m-sample2-U1-latents.csv", delimiter=",")
    assert lat.shape == (3, 20)


def test_blockcg_python():  # macau.pyx:67-91
    announce("api blockcg")
    X = _rand(60, 25, 0.2, 19)
    X.data[:] = 1.0
    rng = np.random.default_rng(20)
    B = np.asfortranarray(rng.standard_normal((3, 25)))
    out = macau_amd.blockcg(X, B, 0.5, tol=1e-10)
    assert out.shape == (2,) and out[0] >= 1


# ---------------------------------------------------------------- values and order
def _oracle_compare(res, ti, tv, ref, nsamples, colnames, tol=1e-4):
    """res.prediction in getTestData order (index columns and y exact), y_pred
    and y_pred_std against the oracle run on the same data and seed."""
    import oracle as O  # noqa: F401  (the checker)
    pred = res.prediction
    assert list(pred.columns) == list(colnames) + ["y", "y_pred", "y_pred_std"]
    np.testing.assert_array_equal(pred[list(colnames)].to_numpy(), ti)
    np.testing.assert_array_equal(pred["y"].to_numpy(), tv)
    yp = pred["y_pred"].to_numpy()
    err = np.linalg.norm(yp - ref.pred) / np.linalg.norm(ref.pred)
    sd_ref = np.sqrt(ref.pred_var / (nsamples - 1))
    err_sd = np.linalg.norm(pred["y_pred_std"].to_numpy() - sd_ref) / np.linalg.norm(sd_ref)
    print("api y_pred rel err %.3e, y_pred_std rel err %.3e, rmse %.6f vs %.6f" % (err, err_sd, res.rmse_test,
                                                                                  ref.rmse))
    assert err < tol and err_sd < 10 * tol
    assert abs(res.rmse_test - ref.rmse) <= tol * ref.rmse


@pytest.mark.parametrize("sampler", ["plain", "batched"])
def test_api_matrix_predictions_match_oracle(sampler):
    """bpmf() on a sparse matrix: prediction rows in the CSC order of Ytest
    (sparsetensor.cpp:63-76), duplicates summed, values equal to the oracle's
    posterior means (bpmfutils.h:86-114) and stds (macau.cpp:175-188)."""
    import oracle as O
    announce("api matrix predictions vs oracle sampler=%s" % sampler)
    rng = np.random.default_rng(31)
    nrow, ncol, K = 400, 250, 32 if sampler == "batched" else 8
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.08)
    y = np.round(rng.standard_normal(len(r)) * 2) / 2 + 3.0
    perm = rng.permutation(len(r))
    te, tr = perm[:900], perm[900:]
    # the test matrix lists cells in shuffled order, one of them twice
    te = np.concatenate([te, te[:1]])
    Ytr = scipy.sparse.coo_matrix((y[tr], (r[tr], c[tr])), shape=(nrow, ncol))
    Yte = scipy.sparse.coo_matrix((y[te], (r[te], c[te])), shape=(nrow, ncol))
    burnin, nsamples = 2, 3
    res = macau_amd.bpmf(Ytr, Ytest=Yte, num_latent=K, precision=4.0, burnin=burnin, nsamples=nsamples,
                         verbose=False, seed=33, sampler=sampler)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, precision=4.0, burnin=burnin, nsamples=nsamples, seed=33)
    # plain FP32: 1e-4; split-BF16 Gram: 1e-3
    _oracle_compare(res, ti, tv, ref, nsamples, ["row", "col"], 1e-4 if sampler == "plain" else 1e-3)


def test_api_tensor_predictions_match_oracle():
    """macau() on a 3-mode DataFrame whose largest mode is mode 1: prediction
    rows in the mode-0 order of SparseMode (sparsetensor.cpp:178-191), values
    equal to the oracle's (eval_rmse_tensor, bpmfutils.cpp:10-73)."""
    import oracle as O
    announce("api tensor predictions vs oracle")
    rng = np.random.default_rng(41)
    dims = (30, 160, 7)
    n = 6000
    idx = np.stack([rng.integers(0, d, n) for d in dims], 1)
    idx = np.unique(idx, axis=0)
    idx = idx[rng.permutation(len(idx))]
    vals = rng.standard_normal(len(idx))
    te, tr = np.arange(700), np.arange(700, len(idx))
    df = pd.DataFrame(idx, columns=["A", "B", "C"])
    df["value"] = vals
    Ytr, Yte = df.iloc[tr], df.iloc[te]
    # the training frame must span every index so both sides see the same dims
    assert all(Ytr[c].max() + 1 == d for c, d in zip("ABC", dims))
    K, burnin, nsamples = 6, 2, 3
    res = macau_amd.bpmf(Ytr, Ytest=Yte, num_latent=K, precision=3.0, burnin=burnin, nsamples=nsamples,
                         verbose=False, seed=43)
    modes, mean, nnz = O.tensor_rows(idx[tr], vals[tr], list(dims))
    ti, tv = O.tensor_test(idx[te], vals[te], list(dims))
    ref = O.run(modes, mean, nnz, list(dims), K, ti, tv, is_tensor=True, precision=3.0, burnin=burnin,
                nsamples=nsamples, seed=43)
    _oracle_compare(res, ti, tv, ref, nsamples, ["A", "B", "C"])


def test_save_model_contents(tmp_path):
    """Model files (macau.cpp:191-205, latentprior.cpp:307-315) hold the
    state the C ABI reports, written as writeToCSVfile does (bpmfutils.h:169-173:
    6 significant digits, "," between columns, "\\n" between rows)."""
    from _gpu_helpers import GpuModel
    from macau_amd import _lib
    from macau_amd._lib import check, ptr
    announce("api save_prefix file contents")
    rng = np.random.default_rng(51)
    nrow, ncol, K, nfeat = 40, 30, 5, 12
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.3)
    y = rng.standard_normal(len(r)) + 2.0
    fr, fc = np.nonzero(rng.random((nrow, nfeat)) < 0.3)
    m = GpuModel(2, K, 3.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol])
    m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6)
    m.prior()
    L = _lib.lib()
    prefix = str(tmp_path / "s")
    check(L.macau_set_samples(m.h, 1, 2))
    check(L.macau_set_seed(m.h, 52))
    check(L.macau_set_verbose(m.h, 0))
    check(L.macau_set_save_model(m.h, 1))
    check(L.macau_set_save_prefix(m.h, prefix.encode()))
    check(L.macau_run(m.h))

    def text(name):
        with open(prefix + name) as f:
            return f.read()

    def expect(mat):
        return "\n".join(",".join("%.6g" % v for v in row) for row in np.atleast_2d(mat))

    assert text("-meanvalue.csv") == "%.6g" % L.macau_get_mean_value(m.h)
    assert abs(L.macau_get_mean_value(m.h) - y.mean()) < 1e-12
    for mode, n in ((0, nrow), (1, ncol)):
        lat = m.latents(mode, n)  # (n, K): the file is K rows x n columns
        assert text("-sample2-U%d-latents.csv" % (mode + 1)) == expect(lat.T)
        mu, _ = m.hyper(mode)
        assert text("-sample2-U%d-latentmean.csv" % (mode + 1)) == expect(mu.reshape(K, 1))
    beta = m.link(0, nfeat)  # (nfeat, K): the file is K x nfeat
    assert text("-sample2-U1-link.csv") == expect(beta.T)
    assert not os.path.exists(prefix + "-sample2-U2-link.csv")
    assert os.path.exists(prefix + "-sample1-U1-link.csv")


@pytest.mark.parametrize("nmodes", [7, 10])
def test_api_tensor_many_modes_matches_oracle(nmodes):
    """DataFrames with up to 10 index columns (macau.pyx:283-284): the
    predictions of a short chain equal the oracle's (plain sampler; the batched
    sampler covers up to 5 modes)."""
    import oracle as O
    announce("api tensor with %d modes" % nmodes)
    rng = np.random.default_rng(80 + nmodes)
    dims = [12, 9, 7, 6, 5, 4, 4, 3, 3, 2][:nmodes]
    cols = ["m%d" % m for m in range(nmodes)]
    idx = np.unique(np.stack([rng.integers(0, d, 5000) for d in dims], axis=1), axis=0)
    idx = idx[rng.permutation(len(idx))]
    F = [rng.standard_normal((d, 3)) * 0.9 for d in dims]
    vals = np.prod([F[m][idx[:, m]] for m in range(nmodes)], axis=0).sum(1) + 0.1 * rng.standard_normal(len(idx))
    te, tr = np.arange(400), np.arange(400, len(idx))
    df = pd.DataFrame(idx, columns=cols)
    df["value"] = vals
    Ytr, Yte = df.iloc[tr], df.iloc[te]
    assert all(Ytr[c].max() + 1 == d for c, d in zip(cols, dims))
    K, burnin, nsamples = 4, 2, 3
    res = macau_amd.bpmf(Ytr, Ytest=Yte, num_latent=K, precision=10.0, burnin=burnin, nsamples=nsamples,
                         verbose=False, seed=45)
    modes, mean, nnz = O.tensor_rows(idx[tr], vals[tr], list(dims))
    ti, tv = O.tensor_test(idx[te], vals[te], list(dims))
    ref = O.run(modes, mean, nnz, list(dims), K, ti, tv, is_tensor=True, precision=10.0, burnin=burnin,
                nsamples=nsamples, seed=45)
    _oracle_compare(res, ti, tv, ref, nsamples, cols)
