"""Pins for the oracle's Macau prior update and Gibbs-loop order.  CPU only.

* ``sample_beta`` (latentprior.cpp:272-293) draws
  beta = Ft_y A^{-1}, A = F^T F + lambda_beta I, with
  Ft_y = (U + E1 - mu 1^T) F + sqrt(lambda_beta) E2 and the columns of E1, E2
  i.i.d. N(0, Lambda^{-1}).  E1 F + sqrt(lambda_beta) E2 is matrix-normal
  MN(0, Lambda^{-1}, F^T F + lambda_beta I), so beta is matrix-normal with
  mean (U - mu 1^T) F A^{-1} and covariance Lambda^{-1} (x) A^{-1}: whitened as
  Lambda^{1/2} (beta - mean) A^{1/2} its entries are i.i.d. N(0, 1).  Checked
  over 4000 draws on both solver branches.
* ``update_prior`` (latentprior.cpp:254-264) equals its pinned pieces in the
  reference order: residual U - Uhat, CondNormalWishart with
  WI + lambda_beta beta beta^T and df + nfeat, sample_beta with the new
  (mu, Lambda), Uhat = beta F^T, lambda_beta from the gamma posterior
  (latentprior.cpp:317-331) written out here from its formula.
* One Gibbs iteration of ``run`` equals "sample every mode, then update every
  prior, then the noise, then the evaluation" (macau.cpp:121-132), composed
  here from the pinned per-piece oracle functions.

Each pin is also run against oracle variants built with a planted mistake
(``ORC_MUTANT`` in oracle.c) and must fail on them.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

SEED = 424242
ORACLE_SRC = os.path.join(os.path.dirname(O.__file__), "oracle.c")


def _sqrtm_spd(A):
    d, Q = np.linalg.eigh(A)
    return (Q * np.sqrt(d)) @ Q.T


def _beta_problem():
    rng = np.random.default_rng(3)
    K, N, nfeat = 3, 10, 6
    Fd = np.zeros((N, nfeat))
    mask = rng.random((N, nfeat)) < 0.5
    mask[np.arange(nfeat) % N, np.arange(nfeat)] = True  # every feature present
    Fd[mask] = rng.uniform(0.5, 1.5, mask.sum())
    r, c = np.nonzero(Fd)
    F, Ft = O.csr_pair(r, c, Fd[r, c], N, nfeat)
    U = rng.standard_normal((N, K))
    mu = np.array([1.0, -2.0, 0.5])
    Q = np.linalg.qr(rng.standard_normal((K, K)))[0]
    Lam = (Q * np.array([0.4, 1.5, 3.0])) @ Q.T
    return K, N, nfeat, Fd, F, Ft, U, mu, Lam


def _check_beta_moments(branch, ndraw=4000):
    K, N, nfeat, Fd, F, Ft, U, mu, Lam = _beta_problem()
    lb = 5.0
    FtF = O.at_mul_a(F, Ft) if branch == "ftf" else None
    A = Fd.T @ Fd + lb * np.eye(nfeat)
    mean = np.linalg.solve(A, Fd.T @ (U - mu))  # (nfeat, K) = ((U - mu) F A^{-1})^T
    Ah, Lh = _sqrtm_spd(A), _sqrtm_spd(Lam)
    W = np.empty((ndraw, nfeat * K))
    for t in range(ndraw):
        beta, _ = O.sample_beta(U, mu, Lam, F, Ft, lb, SEED, t, 0, FtF=FtF, tol=1e-12)
        W[t] = (Ah @ (beta - mean) @ Lh).ravel()
    wm = W.mean(0)
    C = W.T @ W / ndraw
    assert np.abs(wm).max() < 4.5 / np.sqrt(ndraw), "E[beta] differs from (U - mu) F A^-1"
    assert np.abs(C - np.eye(nfeat * K)).max() < 0.1, "Cov[beta] differs from Lambda^-1 (x) A^-1"


def _check_update_prior_composition(branch):
    rng = np.random.default_rng(5)
    K, N, nfeat, Fd, F, Ft, U, _, _ = _beta_problem()
    FtF = O.at_mul_a(F, Ft) if branch == "ftf" else None
    beta = 0.3 * rng.standard_normal((nfeat, K))
    Uhat = Fd @ beta
    lb = 2.5
    for it in (0, 1):
        got = O.macau_update_prior(U, F, Ft, beta, Uhat, lb, SEED, it, mode=1, FtF=FtF, tol=1e-12)
        # the reference's order, from the pinned pieces
        R = U - Uhat
        T = np.eye(K) + lb * (beta.T @ beta)
        mu, Lam = O.cond_normal_wishart(R, np.zeros(K), 2.0, T, K + nfeat, SEED, it, 1)
        beta_n, _ = O.sample_beta(U, mu, Lam, F, Ft, lb, SEED, it, 1, FtF=FtF, tol=1e-12)
        Uhat_n = Fd @ beta_n
        nux = 1e-3 + K * nfeat
        mux = nux / (1e-3 + np.trace((beta_n.T @ beta_n) @ Lam))
        lb_n = O.gamma(SEED, nux / 2.0, 0, it, O.tag(O.TAG_LAMBDA_BETA, 1)) * (2.0 * mux / nux)
        np.testing.assert_allclose(got[0], mu, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(got[1], Lam, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(got[2], beta_n, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(got[3], Uhat_n, rtol=1e-8, atol=1e-10)
        assert got[4] == pytest.approx(lb_n, rel=1e-8)
        beta, Uhat, lb = got[2], got[3], got[4]


def _loop_problem(with_side):
    rng = np.random.default_rng(11)
    nrow, ncol, K, nfeat = 30, 20, 3, 12
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.35)
    y = rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:25], perm[25:]
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = None
    if with_side:
        Fd = (rng.random((nrow, nfeat)) < 0.3) * rng.uniform(0.5, 1.5, (nrow, nfeat))
        fr, fc = np.nonzero(Fd)
        side = [O.csr_pair(fr, fc, Fd[fr, fc], nrow, nfeat), None]
    return modes, mean, nnz, [nrow, ncol], K, ti, tv, side


def _check_loop_order(with_side, iters=3):
    modes, mean, nnz, dims, K, ti, tv, side = _loop_problem(with_side)
    burnin = 1
    res = O.run(modes, mean, nnz, dims, K, ti, tv, adaptive=True, side=side, burnin=burnin, nsamples=5, seed=SEED,
                max_iters=iters, want_U=True, tol=1e-12)
    # MacauX::init and the priors' init (macau.cpp:72-86, latentprior.cpp:84-99, 181-216)
    U = [np.zeros((d, K)) for d in dims]
    mu = [np.zeros(K) for _ in dims]
    Lam = [10.0 * np.eye(K) for _ in dims]
    beta = np.zeros((side[0][0].ncol, K)) if side else None
    Uhat = np.zeros((dims[0], K))
    lb = 5.0
    var = np.mean((modes[0].vals - mean) ** 2)  # AdaptiveGaussianNoise::init (noisemodels.cpp:35-53)
    alpha, alpha_max = 2.0 / var, 11.0 / var
    pred, pvar = np.zeros(len(tv)), np.zeros(len(tv))
    for i in range(iters):
        for m in range(2):  # sample every mode ...
            mr = (mu[m] + Uhat) if (side and m == 0) else None
            U[m] = O.sample_latents(K, modes[m], mean, [U[1 - m]], alpha, mu[m], Lam[m], SEED, i, m, mu_rows=mr)
        for m in range(2):  # ... then update every prior ...
            if side and m == 0:
                mu[m], Lam[m], beta, Uhat, lb, _ = O.macau_update_prior(U[m], side[0][0], side[0][1], beta, Uhat, lb,
                                                                         SEED, i, mode=m, tol=1e-12)
            else:
                mu[m], Lam[m] = O.cond_normal_wishart(U[m], np.zeros(K), 2.0, np.eye(K), K, SEED, i, m)
        sse = O.train_sse(K, modes[0], mean, U)  # ... then the noise ...
        alpha = O.noise_alpha(sse, nnz, var, False, alpha_max, SEED, i)
        r1, ravg = O.eval_rmse(K, ti, tv, mean, U, max(0, i - burnin), pred, pvar)  # ... then the evaluation
        assert res.alpha_trace[i] == pytest.approx(alpha, rel=1e-10)
        assert res.rmse_trace[i] == pytest.approx(ravg, rel=1e-10)
    for m in range(2):
        np.testing.assert_allclose(res.U[m], U[m], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(res.pred, pred, rtol=1e-9, atol=1e-12)
    if side:
        assert res.lambda_beta[0] == pytest.approx(lb, rel=1e-9)


@pytest.mark.parametrize("branch", ["cg", "ftf"])
def test_sample_beta_matrix_normal_moments(branch):
    _check_beta_moments(branch)


def test_sample_beta_branches_agree_per_draw():
    """The block CG (tol 1e-12) and the F^T F Cholesky solve give the same
    beta for the same draws."""
    K, N, nfeat, Fd, F, Ft, U, mu, Lam = _beta_problem()
    FtF = O.at_mul_a(F, Ft)
    for t in range(20):
        a, _ = O.sample_beta(U, mu, Lam, F, Ft, 5.0, SEED, t, 0, FtF=FtF)
        b, its = O.sample_beta(U, mu, Lam, F, Ft, 5.0, SEED, t, 0, tol=1e-12)
        assert its > 0
        np.testing.assert_allclose(b, a, rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("branch", ["cg", "ftf"])
def test_macau_update_prior_composition(branch):
    _check_update_prior_composition(branch)


@pytest.mark.parametrize("with_side", [False, True])
def test_gibbs_iteration_order(with_side):
    _check_loop_order(with_side)


# ------------------------------------------------------------------ planted mistakes
MUTANTS = [
    (1, "sqrt(lambda_beta) dropped from the feature noise", lambda: _check_beta_moments("ftf", 1500)),
    (2, "mu not subtracted from U", lambda: _check_beta_moments("ftf", 1500)),
    (3, "beta noise drawn with covariance Lambda instead of Lambda^-1", lambda: _check_beta_moments("ftf", 1500)),
    (4, "WI + beta beta^T without lambda_beta", lambda: _check_update_prior_composition("ftf")),
    (5, "priors updated before the modes are sampled", lambda: _check_loop_order(False)),
    (6, "Normal-Wishart on U instead of the residual U - Uhat", lambda: _check_update_prior_composition("ftf")),
    (7, "noise precision updated between the modes", lambda: _check_loop_order(False)),
]
# Updating each mode's prior right after that mode is sampled is not a
# mistake the chain can show: mode m's prior enters only mode m's own draw,
# and its update reads only U_m, so that order gives the same chain.


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


@pytest.mark.parametrize("n,what,check", MUTANTS, ids=[str(m[0]) for m in MUTANTS])
def test_pins_fail_on_planted_mistake(n, what, check, mutant_dir):
    path = str(mutant_dir / ("liboracle_m%d.so" % n))
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-fopenmp", "-D_GNU_SOURCE", "-DORC_MUTANT=%d" % n,
                    "-shared", "-o", path, ORACLE_SRC, "-lm"], check=True)
    check()  # the correct oracle passes ...
    with O.use_library(path):
        with pytest.raises(AssertionError):
            check()  # ... and the planted mistake is caught
