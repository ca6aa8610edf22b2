"""Pins for the oracle's univariate prior (MacauOnePrior, macauoneprior.cpp):
the reference's known answers for the two sparse primitives of sample_beta
(tests.cpp:338-403), the coordinate-wise row update against a Gauss-Seidel
sweep written in matrix form, the stationary laws of the row sampler and of
the feature sweep (closed-form Gaussian posteriors), the per-latent
lambda_beta gamma law, the composition of update_prior, and the reference's
Python quality bound for univariate tensor factorisation
(test_macau.py:191-207).  CPU only."""
import json
import math
import os

import numpy as np
import pytest
import scipy.stats

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 777


@pytest.fixture(scope="module")
def kat():
    with open(os.path.join(HERE, "golden", "reference_kat.json")) as f:
        return json.load(f)


def _feat(kat, binary):
    f = kat["feat"]
    vals = np.ones(len(f["vals"])) if binary else np.array(f["vals"])
    return O.csr_pair(f["rows"], f["cols"], vals, f["nrow"], f["ncol"])


@pytest.mark.parametrize("binary", [True, False])
def test_at_mul_bt_kat(kat, binary):  # tests.cpp:338-368
    a = kat["At_mul_Bt"]
    _, Ft = _feat(kat, binary)
    y = O.at_mul_bt_col(Ft, a["col"], np.array(a["B"]).T)
    np.testing.assert_allclose(y, a["Y_bin"] if binary else a["Y_csr"], rtol=1e-12)


@pytest.mark.parametrize("binary", [True, False])
def test_add_acol_mul_bt_kat(kat, binary):  # tests.cpp:370-403
    a = kat["add_Acol_mul_bt"]
    _, Ft = _feat(kat, binary)
    Z = O.add_acol_mul_bt(Ft, a["col"], a["b"], np.array(a["Z"]).T)
    np.testing.assert_allclose(Z.T, a["Z_bin"] if binary else a["Z_csr"], rtol=1e-12, atol=1e-12)


def _gauss_seidel(K, rows, mean, others, alpha, lam, mu_rows, U, seed, it, mode):
    """One Gauss-Seidel sweep of each row's conditional N(P^{-1} h, P^{-1}),
    P = diag(lam) + alpha sum c c^T, h = lam * m_n + alpha sum (y - mean) c,
    with the latent normals: u_d <- (h_d - sum_{j != d} P_dj u_j) / P_dd
    + z_d / sqrt(P_dd)."""
    out = np.array(U, dtype=float)
    for n in range(rows.nrows):
        P = np.diag(lam).astype(float)
        h = lam * mu_rows[n]
        for e in range(rows.row_ptr[n], rows.row_ptr[n + 1]):
            c = np.ones(K)
            for t, F in enumerate(others):
                c = c * F[rows.idx[e, t]]
            P += alpha * np.outer(c, c)
            h = h + alpha * (rows.vals[e] - mean) * c
        z = O.normals(SEED if seed is None else seed, n, K, it, O.tag(O.TAG_LATENT, mode))
        u = out[n]
        for d in range(K):
            s = h[d] - P[d] @ u + P[d, d] * u[d]
            u[d] = s / P[d, d] + z[d] / math.sqrt(P[d, d])
    return out


@pytest.mark.parametrize("tensor", [False, True])
def test_sample_latents_uni_is_gauss_seidel(tensor):
    rng = np.random.default_rng(1)
    K = 5
    if tensor:
        dims = [40, 12, 6]
        idx = np.stack([rng.integers(0, d, 400) for d in dims], 1)
        y = rng.standard_normal(400)
        modes, mean, _ = O.tensor_rows(idx, y, dims)
        others = [rng.standard_normal((dims[1], K)), rng.standard_normal((dims[2], K))]
    else:
        r, c = np.nonzero(rng.random((40, 25)) < 0.2)
        y = rng.standard_normal(len(r))
        modes, mean, _ = O.matrix_rows(r, c, y, 40, 25)
        others = [rng.standard_normal((25, K))]
    rows = modes[0]
    lam = rng.random(K) + 0.5
    mu_rows = rng.standard_normal((rows.nrows, K)) * 0.3
    U = rng.standard_normal((rows.nrows, K))
    got = O.sample_latents_uni(K, rows, mean, others, 2.5, lam, mu_rows, U, SEED, 4, 0)
    want = _gauss_seidel(K, rows, mean, others, 2.5, lam, mu_rows, U, SEED, 4, 0)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)


def test_sample_latents_uni_stationary_law():
    """Repeated sweeps of one row converge to N(P^{-1} h, P^{-1})."""
    rng = np.random.default_rng(2)
    K, ncol = 3, 12
    c = np.arange(ncol)
    y = rng.standard_normal(ncol)
    modes, mean, _ = O.matrix_rows(np.zeros(ncol, int), c, y, 1, ncol)
    V = rng.standard_normal((ncol, K)) * 0.4
    lam = np.array([1.0, 2.0, 0.7])
    m = np.array([[0.2, -0.1, 0.3]])
    alpha = 1.5
    P = np.diag(lam) + alpha * V.T @ V
    h = lam * m[0] + alpha * V.T @ (y - mean)
    S = np.linalg.inv(P)
    u = np.zeros((1, K))
    draws = []
    for t in range(20000):
        u = O.sample_latents_uni(K, modes[0], mean, [V], alpha, lam, m, u, SEED, t, 0)
        draws.append(u[0].copy())
    D = np.array(draws[100:])
    np.testing.assert_allclose(D.mean(0), S @ h, atol=0.03)
    np.testing.assert_allclose(np.cov(D.T), S, atol=0.03, rtol=0.06)


def test_beta_sweep_stationary_law():
    """Feature sweeps with U, mu, lambda, lambda_beta fixed converge, for each
    latent d, to N(S_d lambda_d F^T (U_d - mu_d), S_d) with
    S_d = (lambda_d F^T F + lambda_beta_d I)^{-1}."""
    rng = np.random.default_rng(3)
    N, nfeat, K = 30, 4, 2
    Fd = (rng.random((N, nfeat)) < 0.5) * rng.standard_normal((N, nfeat))
    fr, fc = np.nonzero(Fd)
    F, Ft = O.csr_pair(fr, fc, Fd[fr, fc], N, nfeat)
    colsq = (Fd ** 2).sum(0)
    U = rng.standard_normal((N, K))
    mu = np.array([0.1, -0.2])
    lam = np.array([1.5, 0.6])
    lb = np.array([2.0, 5.0])
    beta = np.zeros((nfeat, K))
    draws = []
    for t in range(12000):
        beta = O.beta_sweep_uni(Ft, colsq, lam, lb, mu, U, Fd @ beta, beta, SEED, t, 0)
        draws.append(beta.copy())
    D = np.array(draws[100:])
    for d in range(K):
        S = np.linalg.inv(lam[d] * Fd.T @ Fd + lb[d] * np.eye(nfeat))
        m = S @ (lam[d] * Fd.T @ (U[:, d] - mu[d]))
        np.testing.assert_allclose(D[:, :, d].mean(0), m, atol=0.02)
        np.testing.assert_allclose(np.cov(D[:, :, d].T), S, atol=0.02, rtol=0.08)


def test_lambda_beta_uni_law():
    """lambda_beta_d (0.1 + |beta_d|^2 / 2) ~ Gamma(0.1 + nfeat / 2)."""
    rng = np.random.default_rng(4)
    beta = rng.standard_normal((7, 3))
    a = 0.1 + 7 / 2.0
    b = 0.1 + (beta ** 2).sum(0) / 2
    x = np.array([O.lambda_beta_uni(beta, SEED, t, 1) * b for t in range(4000)])
    for d in range(3):
        assert scipy.stats.kstest(x[:, d], scipy.stats.gamma(a).cdf).pvalue > 1e-3


def test_update_prior_composition():
    """update_prior = Normal-Wishart of U - Uhat (diagonal kept), the feature
    sweep from the previous Uhat, compute_uhat, then lambda_beta
    (macauoneprior.cpp:189-196)."""
    rng = np.random.default_rng(5)
    N, nfeat, K = 50, 6, 3
    Fd = (rng.random((N, nfeat)) < 0.4).astype(float)
    fr, fc = np.nonzero(Fd)
    F, Ft = O.csr_pair(fr, fc, np.ones(len(fr)), N, nfeat)
    colsq = Fd.sum(0)
    U = rng.standard_normal((N, K))
    beta0 = rng.standard_normal((nfeat, K)) * 0.3
    Uhat0 = Fd @ beta0
    lb0 = np.array([3.0, 4.0, 5.0])
    mu, lam, beta, Uhat, lb = O.macau_one_update_prior(U, F, Ft, colsq, np.zeros(K), np.full(K, 10.0), beta0, Uhat0,
                                                       lb0, SEED, 6, 0)
    mu_w, Lam = O.cond_normal_wishart(U - Uhat0, np.zeros(K), 2.0, np.eye(K), K, SEED, 6, 0)
    np.testing.assert_allclose(mu, mu_w, rtol=1e-12)
    np.testing.assert_allclose(lam, np.diag(Lam), rtol=1e-12)
    beta_w = O.beta_sweep_uni(Ft, colsq, np.diag(Lam), lb0, mu_w, U, Uhat0, beta0, SEED, 6, 0)
    np.testing.assert_allclose(beta, beta_w, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(Uhat, Fd @ beta_w, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(lb, O.lambda_beta_uni(beta_w, SEED, 6, 0), rtol=1e-12)


def test_run_tensor_univariate_quality():  # test_macau.py:191-207
    rng = np.random.default_rng(6)
    A, B, C = rng.standard_normal((30, 2)), rng.standard_normal((4, 2)), rng.standard_normal((2, 2))
    idx = np.array([(i, j, k) for i in range(30) for j in range(4) for k in range(2)])
    val = np.sum(A[idx[:, 0]] * B[idx[:, 1]] * C[idx[:, 2]], 1)
    perm = rng.permutation(len(idx))
    te, tr = perm[:len(idx) // 5], perm[len(idx) // 5:]
    dims = [30, 4, 2]
    modes, mean, nnz = O.tensor_rows(idx[tr], val[tr], dims)
    ti, tv = O.tensor_test(idx[te], val[te], dims)
    ar, ac = np.nonzero(A != 0)
    side = [O.csr_pair(ar, ac, A[ar, ac], 30, 2), None, None]
    res = O.run(modes, mean, nnz, dims, 4, ti, tv, is_tensor=True, precision=50.0, side=side, burnin=20,
                nsamples=20, seed=SEED, uni=[True, False, False])
    assert res.rmse < 0.5


def test_run_matrix_univariate_recovers():
    """Univariate Macau on a matrix with informative row features: test RMSE
    close to the multivariate Macau run on the same data."""
    rng = np.random.default_rng(7)
    nrow, ncol, nfeat, K = 150, 60, 40, 6
    Fd = (rng.random((nrow, nfeat)) < 0.1).astype(float)
    fr, fc = np.nonzero(Fd)
    Ut = Fd @ rng.standard_normal((nfeat, 3)) / 2
    Vt = rng.standard_normal((ncol, 3))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.3)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    kw = dict(precision=10.0, side=side, burnin=30, nsamples=30, seed=SEED)
    uni = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, uni=[True, False], **kw)
    multi = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    sd = float(np.std(y))
    assert uni.rmse < 0.6 * sd
    assert uni.rmse < 1.25 * multi.rmse
    assert uni.lambda_beta[0] > 0.0


def test_run_univariate_probit_rejected():
    """MacauOnePrior::sample_latents(ProbitNoise&, ...) throws "Not
    implemented!" (macauoneprior.cpp:260-264)."""
    rng = np.random.default_rng(8)
    r, c = np.nonzero(rng.random((20, 10)) < 0.4)
    y = (rng.random(len(r)) < 0.5).astype(float)
    modes, mean, nnz = O.matrix_rows(r, c, y, 20, 10)
    ti, tv = O.matrix_test(r[:3], c[:3], y[:3], 20, 10)
    side = [O.csr_pair([0, 1], [0, 1], [1.0, 1.0], 20, 2), None]
    with pytest.raises(RuntimeError):
        O.run(modes, mean, nnz, [20, 10], 3, ti, tv, probit=True, side=side, uni=[True, False], burnin=1,
              nsamples=1, seed=SEED)
