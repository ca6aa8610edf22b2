"""Pins for oracle functions the reference's tests leave open: textbook
library routines (scipy Cholesky / triangular solves / dense solves),
distributional checks of every sampled quantity, and end-to-end quality
bounds from the reference's Python tests.  CPU only."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse
import scipy.stats

import oracle as O

SEED = 12345


def _random_rows(rng, nrows, ncols, density):
    mask = rng.random((nrows, ncols)) < density
    r, c = np.nonzero(mask)
    y = rng.standard_normal(len(r))
    return r, c, y


def _textbook_rows(K, rows, mean, others, alpha, mu, Lam, it, mode, mu_rows=None):
    """u_n = L^{-T}(L^{-1} b + z_n) with scipy's Cholesky and triangular solves."""
    U = np.zeros((rows.nrows, K))
    for n in range(rows.nrows):
        P = Lam.copy()
        b = np.zeros(K)
        for e in range(rows.row_ptr[n], rows.row_ptr[n + 1]):
            c = np.ones(K)
            for t, f in enumerate(others):
                c = c * f[rows.idx[e, t]]
            P += alpha * np.outer(c, c)
            b += alpha * c * (rows.vals[e] - mean)
        m = mu if mu_rows is None else mu_rows[n]
        b += Lam @ m
        L = scipy.linalg.cholesky(P, lower=True)
        z = O.normals(SEED, n, K, it, O.tag(O.TAG_LATENT, mode))
        U[n] = scipy.linalg.solve_triangular(L.T, scipy.linalg.solve_triangular(L, b, lower=True) + z, lower=False)
    return U


def _spd(rng, K, scale=1.0):
    A = rng.standard_normal((K, K))
    return scale * (A @ A.T / K + np.eye(K))


@pytest.mark.parametrize("K", [1, 3, 10, 17])
def test_matrix_sampler_matches_textbook(K):
    rng = np.random.default_rng(K)
    r, c, y = _random_rows(rng, 40, 25, 0.3)
    modes, mean, _ = O.matrix_rows(r, c, y, 40, 25)
    V = rng.standard_normal((25, K))
    Lam, mu = _spd(rng, K, 2.0), rng.standard_normal(K)
    U = O.sample_latents(K, modes[0], mean, [V], 3.0, mu, Lam, SEED, 7, 0)
    ref = _textbook_rows(K, modes[0], mean, [V], 3.0, mu, Lam, 7, 0)
    np.testing.assert_allclose(U, ref, rtol=1e-10, atol=1e-12)


def test_tensor_sampler_matches_textbook():
    rng = np.random.default_rng(5)
    K, dims = 6, [9, 7, 4]
    idx = np.stack([rng.integers(0, d, 60) for d in dims], axis=1)
    vals = rng.standard_normal(60)
    modes, mean, _ = O.tensor_rows(idx, vals, dims)
    fs = [rng.standard_normal((d, K)) for d in dims]
    Lam, mu = _spd(rng, K), rng.standard_normal(K)
    for mode in range(3):
        others = [fs[q] for q in range(3) if q != mode]
        U = O.sample_latents(K, modes[mode], mean, others, 5.0, mu, Lam, SEED, 2, mode)
        ref = _textbook_rows(K, modes[mode], mean, others, 5.0, mu, Lam, 2, mode)
        np.testing.assert_allclose(U, ref, rtol=1e-10, atol=1e-12)


def test_sampler_per_row_mean_matches_textbook():
    rng = np.random.default_rng(9)
    K = 5
    r, c, y = _random_rows(rng, 30, 12, 0.4)
    modes, mean, _ = O.matrix_rows(r, c, y, 30, 12)
    V = rng.standard_normal((12, K))
    Lam, mu = _spd(rng, K), rng.standard_normal(K)
    mu_rows = mu + rng.standard_normal((30, K))
    U = O.sample_latents(K, modes[0], mean, [V], 2.0, mu, Lam, SEED, 1, 0, mu_rows=mu_rows)
    ref = _textbook_rows(K, modes[0], mean, [V], 2.0, mu, Lam, 1, 0, mu_rows=mu_rows)
    np.testing.assert_allclose(U, ref, rtol=1e-10, atol=1e-12)


def test_sampler_moments():
    """E[u] = P^{-1} b and Cov[u] = P^{-1} over independent iterations."""
    rng = np.random.default_rng(11)
    K = 3
    r = np.zeros(6, dtype=int)
    c = np.arange(6)
    y = rng.standard_normal(6) + 1.0
    modes, mean, _ = O.matrix_rows(r, c, y, 1, 6)
    V = rng.standard_normal((6, K))
    Lam, mu, alpha = _spd(rng, K), rng.standard_normal(K), 1.5
    draws = np.array([O.sample_latents(K, modes[0], mean, [V], alpha, mu, Lam, SEED, it, 0)[0]
                      for it in range(4000)])
    P = Lam + alpha * V.T @ V
    b = alpha * V.T @ (y - mean) + Lam @ mu
    Pinv = np.linalg.inv(P)
    se = np.sqrt(np.diag(Pinv) / 4000)
    assert np.all(np.abs(draws.mean(0) - Pinv @ b) < 4.5 * se)
    np.testing.assert_allclose(np.cov(draws.T), Pinv, atol=0.08 * np.max(np.abs(Pinv)))


def test_normals_distribution():
    z = np.concatenate([O.normals(SEED, c0, 4000, 3, O.tag(O.TAG_LATENT, 1)) for c0 in range(10)])
    assert scipy.stats.kstest(z, "norm").pvalue > 1e-3
    # distinct streams are uncorrelated
    a = O.normals(SEED, 0, 20000, 0, O.tag(O.TAG_LATENT, 0))
    b = O.normals(SEED, 0, 20000, 0, O.tag(O.TAG_LATENT, 1))
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.03


@pytest.mark.parametrize("shape", [0.3, 1.0, 2.5, 50.0, 4.5e5])
def test_gamma_distribution(shape):
    g = np.array([O.gamma(SEED, shape, c0, 0, O.tag(O.TAG_NOISE)) for c0 in range(3000)])
    assert np.all(g > 0)
    assert scipy.stats.kstest(g, scipy.stats.gamma(shape).cdf).pvalue > 1e-3


def _nw_closed_form(U, mu0, kappa, T, nu):
    N, K = U.shape
    s = U.sum(0)
    S = U.T @ U
    mu_c = (kappa * mu0 + s) / (kappa + N)
    kappa_c = kappa + N
    X = T + S + kappa * np.outer(mu0, mu0) - kappa_c * np.outer(mu_c, mu_c)
    return mu_c, kappa_c, np.linalg.inv(X), nu + N


def test_normal_wishart_distribution():
    """Lambda ~ Wishart(nu_c, T_c): E[Lambda] = nu_c T_c, Lambda_ii / T_c,ii ~ chi2(nu_c);
    mu | Lambda ~ N(mu_c, (kappa_c Lambda)^{-1})."""
    rng = np.random.default_rng(2)
    K, N = 4, 30
    U = rng.standard_normal((N, K)) * 0.7 + 0.2
    mu0, T = np.zeros(K), np.eye(K)
    mu_c, kappa_c, Tc, nu_c = _nw_closed_form(U, mu0, 2.0, T, K)
    lam, q = [], []
    for it in range(3000):
        mu, L = O.cond_normal_wishart(U, mu0, 2.0, T, K, SEED, it, 0)
        lam.append(L)
        d = mu - mu_c
        q.append(kappa_c * d @ L @ d)
    lam = np.array(lam)
    np.testing.assert_allclose(lam.mean(0), nu_c * Tc, atol=0.06 * np.max(np.abs(nu_c * Tc)))
    for i in range(K):
        assert scipy.stats.kstest(lam[:, i, i] / Tc[i, i], scipy.stats.chi2(nu_c).cdf).pvalue > 1e-3
    assert scipy.stats.kstest(np.array(q), scipy.stats.chi2(K).cdf).pvalue > 1e-3


def test_normal_wishart_macau_scale():
    """Non-identity T and large nu (Macau: WI + lambda BB', df + nfeat)."""
    rng = np.random.default_rng(8)
    K, N = 3, 50
    U = rng.standard_normal((N, K))
    T = _spd(rng, K, 0.5)
    mu_c, kappa_c, Tc, nu_c = _nw_closed_form(U, np.zeros(K), 2.0, T, K + 200)
    lam = np.array([O.cond_normal_wishart(U, np.zeros(K), 2.0, T, K + 200, SEED, it, 1)[1] for it in range(1500)])
    np.testing.assert_allclose(lam.mean(0), nu_c * Tc, rtol=0.03, atol=0.03 * np.max(np.abs(nu_c * Tc)))


def test_noise_alpha_distribution():
    """alpha * bN ~ Gamma(aN) for the adaptive noise update (noisemodels.cpp:55-81)."""
    sse, nnz, var = 120.0, 400, 2.0
    aN, bN = 0.5 + nnz / 2, 0.5 * var + sse / 2
    a = np.array([O.noise_alpha(sse, nnz, var, False, 1e9, SEED, it) for it in range(3000)])
    assert scipy.stats.kstest(a * bN, scipy.stats.gamma(aN).cdf).pvalue > 1e-3
    # tensor variant uses b0 = 0.5; clamp at alpha_max
    at = O.noise_alpha(sse, nnz, var, True, 1e9, SEED, 0)
    g = O.gamma(SEED, aN, 0, 0, O.tag(O.TAG_NOISE))
    assert at == pytest.approx(g / (0.5 + sse / 2), rel=1e-14)
    assert O.noise_alpha(sse, nnz, var, False, 0.01, SEED, 0) == 0.01


def test_train_sse_bruteforce():
    rng = np.random.default_rng(4)
    K = 4
    r, c, y = _random_rows(rng, 20, 15, 0.3)
    modes, mean, _ = O.matrix_rows(r, c, y, 20, 15)
    U, V = rng.standard_normal((20, K)), rng.standard_normal((15, K))
    sse = O.train_sse(K, modes[0], mean, [U, V])
    rr, cc, yy = O.sum_duplicates(r, c, y, 20, 15)
    brute = np.sum((np.sum(U[rr] * V[cc], 1) + mean - yy) ** 2)
    assert sse == pytest.approx(brute, rel=1e-12)


@pytest.mark.parametrize("nrhs,binary", [(1, True), (5, False), (32, True), (40, False)])
def test_blockcg_matches_dense_solve(nrhs, binary):
    rng = np.random.default_rng(nrhs)
    nrow, ncol = 60, 45
    F = scipy.sparse.random(nrow, ncol, density=0.15, random_state=nrhs, format="coo")
    vals = np.ones(F.nnz) if binary else F.data
    Fc, Ftc = O.csr_pair(F.row, F.col, vals, nrow, ncol)
    B = rng.standard_normal((ncol, nrhs))
    reg = 0.8
    X, it = O.blockcg(Fc, Ftc, reg, B, 1e-10)
    Fd = scipy.sparse.coo_matrix((vals, (F.row, F.col)), shape=(nrow, ncol)).toarray()
    ref = np.linalg.solve(Fd.T @ Fd + reg * np.eye(ncol), B)
    np.testing.assert_allclose(X, ref, rtol=1e-7, atol=1e-9)
    assert it < 200


def test_blockcg_blocked_split():
    """linop.h:206-226 block split (32 + remainder) equals one-block solves."""
    rng = np.random.default_rng(1)
    F = scipy.sparse.random(50, 30, density=0.2, random_state=3, format="coo")
    Fc, Ftc = O.csr_pair(F.row, F.col, F.data, 50, 30)
    B = rng.standard_normal((30, 45))
    X, _ = O.blockcg(Fc, Ftc, 1.0, B, 1e-10, blocksize=32, excess=8)
    Fd = F.toarray()
    ref = np.linalg.solve(Fd.T @ Fd + np.eye(30), B)
    np.testing.assert_allclose(X, ref, rtol=1e-7, atol=1e-9)


def test_spmm_matches_scipy():
    rng = np.random.default_rng(6)
    F = scipy.sparse.random(70, 40, density=0.1, random_state=6, format="coo")
    Fc, Ftc = O.csr_pair(F.row, F.col, F.data, 70, 40)
    B = rng.standard_normal((40, 7))
    np.testing.assert_allclose(O.csr_A_mul_Bn(Fc, B), F.tocsr() @ B, rtol=1e-12)
    C = rng.standard_normal((70, 7))
    np.testing.assert_allclose(O.csr_A_mul_Bn(Ftc, C), F.T.tocsr() @ C, rtol=1e-12)


def _lowrank_problem(seed, nrow, ncol, K, density, noise):
    rng = np.random.default_rng(seed)
    U, V = rng.standard_normal((nrow, K)), rng.standard_normal((ncol, K))
    mask = rng.random((nrow, ncol)) < density
    r, c = np.nonzero(mask)
    y = np.sum(U[r] * V[c], 1) + noise * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    nte = len(r) // 5
    te, tr = perm[:nte], perm[nte:]
    return (r[tr], c[tr], y[tr]), (r[te], c[te], y[te])


def test_full_run_quality_bpmf():
    (r, c, y), (rt, ct, yt) = _lowrank_problem(0, 200, 150, 3, 0.25, 0.3)
    modes, mean, nnz = O.matrix_rows(r, c, y, 200, 150)
    ti, tv = O.matrix_test(rt, ct, yt, 200, 150)
    res = O.run(modes, mean, nnz, [200, 150], 5, ti, tv, precision=1 / 0.09, burnin=20, nsamples=30, seed=1)
    assert res.rmse < 0.45


def test_full_run_quality_adaptive():
    (r, c, y), (rt, ct, yt) = _lowrank_problem(1, 150, 120, 2, 0.3, 0.5)
    modes, mean, nnz = O.matrix_rows(r, c, y, 150, 120)
    ti, tv = O.matrix_test(rt, ct, yt, 150, 120)
    res = O.run(modes, mean, nnz, [150, 120], 4, ti, tv, adaptive=True, burnin=20, nsamples=30, seed=2)
    assert res.rmse < 0.7
    # the learned precision approaches 1/0.25 (capped at alpha_max)
    assert 2.0 < res.alpha_trace[-1] < 6.0


def test_full_run_quality_tensor():
    """test_macau.py:132-146: 3-mode toy tensor, precision 50, RMSE < 0.5."""
    rng = np.random.default_rng(7)
    A, B, C = rng.standard_normal((15, 2)), rng.standard_normal((20, 2)), rng.standard_normal((3, 2))
    idx = np.array([(i, j, k) for i in range(15) for j in range(20) for k in range(3)])
    val = np.sum(A[idx[:, 0]] * B[idx[:, 1]] * C[idx[:, 2]], 1)
    perm = rng.permutation(len(idx))
    te, tr = perm[:len(idx) // 5], perm[len(idx) // 5:]
    modes, mean, nnz = O.tensor_rows(idx[tr], val[tr], [15, 20, 3])
    ti, tv = O.tensor_test(idx[te], val[te], [15, 20, 3])
    res = O.run(modes, mean, nnz, [15, 20, 3], 4, ti, tv, is_tensor=True, precision=50.0, burnin=20,
                nsamples=20, seed=3)
    assert res.rmse < 0.5


def test_full_run_quality_macau():
    """Side information that determines the rows' latents improves the fit."""
    rng = np.random.default_rng(10)
    nrow, ncol, K, nfeat = 120, 40, 3, 60
    Fd = (rng.random((nrow, nfeat)) < 0.1).astype(float)
    W = rng.standard_normal((nfeat, K))
    U = Fd @ W / 2.0
    V = rng.standard_normal((ncol, K))
    mask = rng.random((nrow, ncol)) < 0.3
    r, c = np.nonzero(mask)
    y = np.sum(U[r] * V[c], 1) + 0.2 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    fr, fc = np.nonzero(Fd)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    res = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, precision=25.0, side=side, burnin=20, nsamples=20,
                seed=4)
    assert res.rmse < 0.6
    assert res.lambda_beta[0] > 0


@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (45, 32), (130, 40)])
def test_ftf_solve_matches_dense_solve(n, nrhs):
    """F^T F branch of sample_beta: Cholesky of the lower triangle + reg I."""
    rng = np.random.default_rng(n)
    F = scipy.sparse.random(3 * n + 5, n, density=0.2, random_state=n, format="coo")
    A, At = O.csr_pair(F.row, F.col, F.data, F.shape[0], n)
    FtF = O.at_mul_a(A, At)
    Fd = F.toarray()
    np.testing.assert_allclose(FtF, np.tril(Fd.T @ Fd), rtol=1e-13, atol=1e-13)
    B = rng.standard_normal((n, nrhs))
    reg = 0.7
    X = O.ftf_solve(FtF, reg, B)
    np.testing.assert_allclose(X, np.linalg.solve(Fd.T @ Fd + reg * np.eye(n), B), rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("nrow,n", [(200, 60), (40, 90)])
def test_ftf_solve_equals_spectral_solve(nrow, n):
    """The identity the GPU's F^T F branch relies on: with F^T F = Q diag(d) Q^T
    (one eigendecomposition for the whole chain), the oracle's per-iteration
    Cholesky solve of (F^T F + lambda I) X = B equals Q diag(1/(d + lambda))
    Q^T B for every lambda, including a rank-deficient F^T F (more features
    than rows, so d has zeros)."""
    rng = np.random.default_rng(nrow + n)
    F = scipy.sparse.random(nrow, n, density=0.1, random_state=nrow, format="coo")
    F.data[:] = 1.0
    A, At = O.csr_pair(F.row, F.col, np.ones(F.nnz), nrow, n)
    FtF = O.at_mul_a(A, At)
    d, Q = np.linalg.eigh(np.tril(FtF) + np.tril(FtF, -1).T)
    if nrow < n:
        assert np.sum(np.abs(d) < 1e-9 * max(1.0, d.max())) >= n - nrow  # rank-deficient
    B = rng.standard_normal((n, 32))
    for lam in (0.05, 1.0, 5.0, 250.0):
        X = O.ftf_solve(FtF, lam, B)
        Y = Q @ ((Q.T @ B) / (d + lam)[:, None])
        np.testing.assert_allclose(X, Y, rtol=1e-9, atol=1e-10 * np.abs(X).max())


def test_ftf_solve_not_positive_definite():
    FtF = np.array([[1.0, 0.0], [3.0, 1.0]])  # lower triangle of an indefinite matrix
    with pytest.raises(RuntimeError):
        O.ftf_solve(FtF, 0.0, np.ones((2, 1)))


def test_full_run_ftf_branch_matches_cg_branch():
    """Both branches of sample_beta solve the same system; with a tight CG
    tolerance the first iterations agree."""
    rng = np.random.default_rng(12)
    nrow, ncol, K, nfeat = 80, 30, 4, 25
    Fd = (rng.random((nrow, nfeat)) < 0.2).astype(float) * rng.standard_normal((nrow, nfeat))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.3)
    y = rng.standard_normal(len(r))
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(r[:10], c[:10], y[:10], nrow, ncol)
    fr, fc = np.nonzero(Fd)
    side = [O.csr_pair(fr, fc, Fd[fr, fc], nrow, nfeat), None]
    kw = dict(precision=5.0, side=side, burnin=2, nsamples=2, seed=5, max_iters=3, want_U=True, tol=1e-12)
    a = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, use_ftf=[True, False], **kw)
    b = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    assert a.cg_iters_last == 0 and b.cg_iters_last > 0
    np.testing.assert_allclose(a.U[0], b.U[0], rtol=1e-6, atol=1e-8)
    assert a.lambda_beta[0] == pytest.approx(b.lambda_beta[0], rel=1e-6)


def test_sampler_with_six_other_modes_matches_textbook():
    """Rows of a 7-mode tensor: the Hadamard product of six other factors."""
    rng = np.random.default_rng(70)
    K = 5
    dims = [12, 6, 5, 4, 3, 3, 2]
    idx = np.stack([rng.integers(0, d, 400) for d in dims], axis=1)
    vals = rng.standard_normal(400)
    modes, mean, _ = O.tensor_rows(idx, vals, dims)
    fs = [rng.standard_normal((d, K)) * 0.8 for d in dims]
    Lam, mu = _spd(rng, K, 2.0), rng.standard_normal(K)
    for mode in (0, 3, 6):
        others = [fs[q] for q in range(7) if q != mode]
        U = O.sample_latents(K, modes[mode], mean, others, 2.0, mu, Lam, SEED, 5, mode)
        ref = _textbook_rows(K, modes[mode], mean, others, 2.0, mu, Lam, 5, mode)
        np.testing.assert_allclose(U, ref, rtol=1e-10, atol=1e-12)


def test_run_ten_mode_tensor():
    """The reference's Data takes up to 10 index columns (macau.pyx:283-284):
    a 10-mode CP tensor is fitted (test RMSE well below the data's sd)."""
    rng = np.random.default_rng(71)
    dims = [8, 6, 5, 4, 4, 3, 3, 3, 2, 2]
    F = [rng.standard_normal((d, 1)) for d in dims]
    idx = np.stack([rng.integers(0, d, 4000) for d in dims], axis=1)
    idx = np.unique(idx, axis=0)
    val = np.prod([F[m][idx[:, m], 0] for m in range(10)], axis=0)
    perm = rng.permutation(len(idx))
    te, tr = perm[:300], perm[300:]
    modes, mean, nnz = O.tensor_rows(idx[tr], val[tr], dims)
    ti, tv = O.tensor_test(idx[te], val[te], dims)
    res = O.run(modes, mean, nnz, dims, 2, ti, tv, is_tensor=True, precision=50.0, burnin=20, nsamples=20, seed=3)
    assert res.rmse < 0.5 * np.std(val)
