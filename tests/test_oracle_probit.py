"""Pins for the oracle's probit noise model (ProbitNoise, noisemodels.h:88-103;
sample_latent_blas_probit, latentprior.cpp:387-414; truncnorm.cpp;
auc, bpmfutils.h:182-213): the reference's known answers (tests.cpp), scipy
and scikit-learn routines for Phi^{-1}, the truncated normal law and ROC AUC,
brute-force pair counts, and the end-to-end quality bound of the reference's
Python probit test.  CPU only."""
import itertools
import json
import math
import os

import numpy as np
import pytest
import scipy.special
import scipy.stats
from sklearn.metrics import roc_auc_score

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 4242


@pytest.fixture(scope="module")
def kat():
    with open(os.path.join(HERE, "golden", "reference_kat.json")) as f:
        return json.load(f)


def test_norm_cdf_kat(kat):  # tests.cpp:894-899
    for x, want in kat["norm_cdf"]:
        assert O.norm_cdf(x) == pytest.approx(want, rel=1e-14)


def test_inv_norm_cdf_kat(kat):  # tests.cpp:887-892
    assert O.inv_norm_cdf(0.0) == -math.inf
    assert O.inv_norm_cdf(1.0) == math.inf
    for p, want in kat["inv_norm_cdf"]:
        assert O.inv_norm_cdf(p) == pytest.approx(want, rel=1e-13, abs=1e-15)


def test_inv_norm_cdf_matches_ndtri():
    """Against scipy's ndtri from 1e-307 to 1 - 1e-16, both tails."""
    p = np.concatenate([np.logspace(-307, np.log10(0.5), 400), 1.0 - np.logspace(-16, np.log10(0.5), 200)])
    got = np.array([O.inv_norm_cdf(x) for x in p])
    want = scipy.special.ndtri(p)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)


def _rand_truncnorm(mean, std, low, seed, c0):
    """rand_truncnorm(mean, std, low) (truncnorm.cpp:26-31) through the
    oracle's standard form: std (truncnorm0(-abar) + abar) + mean."""
    abar = (low - mean) / std
    return std * (O.truncnorm0(-abar, seed, c0, 0, O.tag(O.TAG_PROBIT)) + abar) + mean


def test_rand_truncnorm_bounds_kat(kat):  # tests.cpp:901-910
    for mean, std, low in kat["rand_truncnorm_bounds"]:
        for c0 in range(10):
            assert _rand_truncnorm(mean, std, low, 1234, c0) >= low


@pytest.mark.parametrize("mean", [-3.0, -1.0, 0.0, 2.5])
def test_truncnorm0_law_icdf(mean):
    """Cut abar = -mean <= 3: inverse-CDF sampling, exactly N(mean, 1)
    truncated to [0, inf) (Kolmogorov-Smirnov and the mean)."""
    n = 20000
    x = np.array([O.truncnorm0(mean, SEED, j, 7, O.tag(O.TAG_PROBIT, 1)) for j in range(n)])
    assert x.min() >= 0.0
    law = scipy.stats.truncnorm(a=-mean, b=np.inf, loc=mean, scale=1.0)
    assert scipy.stats.kstest(x, law.cdf).pvalue > 1e-3
    assert abs(x.mean() - law.mean()) < 5 * law.std() / math.sqrt(n)


@pytest.mark.parametrize("mean", [-8.0, -3.5, -3.0 - 1e-9])
def test_truncnorm0_law_tail(mean):
    """Cut abar = -mean > 3: the reference's tail sampler proposes
    xbar = sqrt(abar^2 - 2 log(1 - u)) and accepts when v <= xbar / abar
    (truncnorm.cpp:14-24).  xbar >= abar makes that test always pass, so
    xbar - abar follows the proposal law P(xbar > x) = exp(-(x^2 - abar^2)/2)
    rather than the truncated normal; the restatement keeps that law."""
    n = 20000
    a = -mean
    x = np.array([O.truncnorm0(mean, SEED, j, 7, O.tag(O.TAG_PROBIT, 1)) for j in range(n)]) + a
    assert x.min() >= a
    assert scipy.stats.kstest(x, lambda t: 1.0 - np.exp(-(np.maximum(t, a) ** 2 - a * a) / 2)).pvalue > 1e-3
    # the first attempt is always accepted: the draw is a closed form of
    # block (c0, 0, it, tag)
    ref = np.sqrt(a * a - 2.0 * np.log1p(-np.array([_u53_first(SEED, j, 7, O.tag(O.TAG_PROBIT, 1))
                                                     for j in range(50)])))
    np.testing.assert_allclose(x[:50], ref, rtol=1e-15)


def _u53_first(seed, c0, it, tg):
    w = O.philox([c0, 0, it, tg], [seed & 0xFFFFFFFF, seed >> 32])
    k = (int(w[0]) >> 6 << 27) | (int(w[1]) >> 5)
    return (k + 0.5) / 2.0 ** 53


def test_truncnorm0_keyed():
    """A draw is a function of (mean, seed, observation, iteration, tag)."""
    t = O.tag(O.TAG_PROBIT, 0)
    a = O.truncnorm0(0.3, SEED, 5, 2, t)
    assert a == O.truncnorm0(0.3, SEED, 5, 2, t)
    assert a != O.truncnorm0(0.3, SEED, 6, 2, t)
    assert a != O.truncnorm0(0.3, SEED, 5, 3, t)
    assert a != O.truncnorm0(0.3, SEED, 5, 2, O.tag(O.TAG_PROBIT, 1))


def test_auc_kat(kat):  # tests.cpp:606-614
    a = kat["auc"]
    assert O.auc(a["pred"], a["test"]) == pytest.approx(a["auc"], rel=1e-12)
    assert math.isnan(O.auc([], []))


def test_auc_matches_roc_auc_and_pair_count():
    rng = np.random.default_rng(3)
    for n in (2, 7, 50, 1000):
        test = (rng.random(n) < 0.4).astype(float)
        test[0], test[-1] = 0.0, 1.0
        pred = rng.standard_normal(n)
        got = O.auc(pred, test)
        assert got == pytest.approx(roc_auc_score(test, pred), rel=1e-12)
        pos, neg = pred[test == 1], pred[test == 0]
        pairs = np.mean(pos[:, None] > neg[None, :])
        assert got == pytest.approx(pairs, rel=1e-12)


def _probit_problem(seed, nrow, ncol, rank, density):
    rng = np.random.default_rng(seed)
    A, B = rng.standard_normal((nrow, rank)), rng.standard_normal((ncol, rank))
    r, c = np.nonzero(rng.random((nrow, ncol)) < density)
    y = (np.sum(A[r] * B[c], 1) > 0).astype(float)
    perm = rng.permutation(len(r))
    nte = len(r) // 5
    te, tr = perm[:nte], perm[nte:]
    return (r[tr], c[tr], y[tr]), (r[te], c[te], y[te])


def test_probit_z_keying_and_sign():
    """z_e = y rand_truncnorm(y u_n.v_e, 1, 0) with the observation's
    mode-order index as key (latentprior.cpp:393-398)."""
    (r, c, y), _ = _probit_problem(1, 40, 30, 3, 0.3)
    modes, mean, nnz = O.matrix_rows(r, c, y, 40, 30)
    rng = np.random.default_rng(2)
    K = 5
    U, V = rng.standard_normal((40, K)), rng.standard_normal((30, K))
    rows = modes[0]
    z = O.probit_z(K, rows, [V], U, SEED, 3, 0)
    t = O.tag(O.TAG_PROBIT, 0)
    for n in range(rows.nrows):
        for e in range(rows.row_ptr[n], rows.row_ptr[n + 1]):
            s = 2.0 * rows.vals[e] - 1.0
            d = float(np.dot(U[n], V[rows.idx[e, 0]]))
            assert z[e] == pytest.approx(s * O.truncnorm0(s * d, SEED, e, 3, t), rel=1e-12)
            assert s * z[e] >= 0.0


def test_eval_probit_running_mean():
    """Phi(u.v) per test cell, Welford mean / sum of squared deviations
    over samples, and the AUC of both (noisemodels.cpp:129-159)."""
    rng = np.random.default_rng(4)
    K, nrow, ncol = 3, 20, 15
    ti = np.stack([rng.integers(0, nrow, 60), rng.integers(0, ncol, 60)], 1).astype(np.int32)
    tv = (rng.random(60) < 0.5).astype(float)
    pred, pvar = np.zeros(60), np.zeros(60)
    seen = []
    for n in range(4):
        U, V = rng.standard_normal((nrow, K)), rng.standard_normal((ncol, K))
        p1, a1, aa = O.eval_probit(K, ti, tv, [U, V], n, pred, pvar)
        want = scipy.stats.norm.cdf(np.sum(U[ti[:, 0]] * V[ti[:, 1]], 1))
        np.testing.assert_allclose(p1, want, rtol=1e-14, atol=1e-300)
        seen.append(want)
        S = np.array(seen)
        np.testing.assert_allclose(pred, S.mean(0), rtol=1e-12)
        np.testing.assert_allclose(pvar, ((S - S.mean(0)) ** 2).sum(0), rtol=1e-10, atol=1e-14)
        assert a1 == pytest.approx(roc_auc_score(tv, p1), rel=1e-12)
        assert aa == pytest.approx(roc_auc_score(tv, pred), rel=1e-12)


def test_probit_run_recovers_signs():
    """Binary low-rank matrix: test AUC well above the reference's 0.55
    gate (test_macau.py:47-62), one-sample AUC too, rmse fields NaN."""
    nrow, ncol = 120, 80
    (r, c, y), (rt, ct, yt) = _probit_problem(5, nrow, ncol, 3, 0.3)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
    res = O.run(modes, mean, nnz, [nrow, ncol], 4, ti, tv, probit=True, burnin=20, nsamples=20, seed=SEED)
    assert res.auc > 0.85 and res.auc1 > 0.75
    assert math.isnan(res.rmse) and math.isnan(res.rmse1)
    assert np.all((res.pred > 0) & (res.pred < 1))
    np.testing.assert_array_equal(res.rmse_trace[-1], res.auc)


def test_probit_run_with_dense_side():  # test_macau.py:47-62
    rng = np.random.default_rng(6)
    A, B = rng.standard_normal((25, 2)), rng.standard_normal((3, 2))
    idx = np.array(list(itertools.product(range(25), range(3))))
    val = (np.sum(A[idx[:, 0]] * B[idx[:, 1]], 1) > 0).astype(float)
    perm = rng.permutation(len(idx))
    te, tr = perm[:15], perm[15:]
    modes, mean, nnz = O.matrix_rows(idx[tr, 0], idx[tr, 1], val[tr], 25, 3)
    ti, tv = O.matrix_test(idx[te, 0], idx[te, 1], val[te], 25, 3)
    fr, fc = np.nonzero(np.ones((25, 2), dtype=bool))
    side = [O.csr_pair(fr, fc, A[fr, fc], 25, 2), None]
    res = O.run(modes, mean, nnz, [25, 3], 4, ti, tv, probit=True, side=side, burnin=20, nsamples=20, seed=SEED,
                use_ftf=[True, False])
    assert res.auc > 0.55


def test_probit_run_rejects_tensor():
    """BPMFPrior::sample_latents(ProbitNoise&, TensorData&) throws
    (latentprior.cpp:112-116)."""
    dims = [6, 5, 4]
    idx = np.array([[0, 0, 0], [1, 2, 3], [5, 4, 1]])
    modes, mean, nnz = O.tensor_rows(idx, np.array([1.0, 0.0, 1.0]), dims)
    ti, tv = O.tensor_test(idx[:1], np.array([1.0]), dims)
    with pytest.raises(RuntimeError):
        O.run(modes, mean, nnz, dims, 3, ti, tv, is_tensor=True, probit=True, burnin=1, nsamples=1, seed=SEED)
