"""GPU parity of the univariate prior (MacauOnePrior, macauoneprior.cpp)
against the FP64 oracle: the coordinate-wise row update (the plain sampler's
Gauss-Seidel variant), the Normal-Wishart diagonal, the feature-sequential
link-matrix sweep (residuals in LDS, and the global-memory form for more
rows than LDS holds), Uhat and the per-latent lambda_beta, on matrices and
a 3-mode tensor, with the other modes on the plain or the batched sampler."""
import numpy as np
import pytest

import oracle as O
from _gpu_helpers import SEED, GpuModel, announce, row_rel_err
from macau_amd import _lib

pytestmark = pytest.mark.gpu

PLAIN = _lib.SAMPLER_PLAIN_FP32
BATCHED = _lib.SAMPLER_BATCHED
TOL_FP32 = 1e-4
TOL_BATCHED = 1e-4


def _side_problem(seed, nrow, ncol, nfeat, density, binary):
    rng = np.random.default_rng(seed)
    Fm = rng.random((nrow, nfeat)) < density
    fr, fc = np.nonzero(Fm)
    fv = np.ones(len(fr)) if binary else rng.standard_normal(len(fr))
    Fd = np.zeros((nrow, nfeat))
    Fd[fr, fc] = fv
    Ut = Fd @ rng.standard_normal((nfeat, 3)) / 2
    Vt = rng.standard_normal((ncol, 3))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.15)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    return (fr, fc, fv), (r[tr], c[tr], y[tr]), (r[te], c[te], y[te])


@pytest.mark.parametrize("K,binary,sampler", [(8, True, "plain"), (8, False, "plain"), (32, True, "batched"),
                                              (70, False, "plain")])
def test_model_univariate_matches_oracle(K, binary, sampler):
    announce("model univariate K=%d binary=%d sampler=%s" % (K, binary, sampler))
    nrow, ncol, nfeat = 300, 120, 50
    (fr, fc, fv), (r, c, y), (rt, ct, yt) = _side_problem(K + binary, nrow, ncol, nfeat, 0.08, binary)
    kind = PLAIN if sampler == "plain" else BATCHED
    tol_cols = TOL_FP32 if sampler == "plain" else TOL_BATCHED
    m = GpuModel(2, K, 10.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([rt, ct], 1), yt)
    m.one_prior((fr, fc, None if binary else fv, nrow, nfeat), lambda_beta=5.0)
    m.prior()
    m.setup(4, 6, SEED, kind, graph=False)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
    side = [O.csr_pair(fr, fc, fv, nrow, nfeat), None]
    kw = dict(precision=10.0, side=side, lambda_beta=5.0, burnin=4, nsamples=6, seed=SEED, uni=[True, False])
    for n_it in (1, 2):
        m.iterate(1)
        ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=n_it, want_U=True, **kw)
        e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
        lb = _lib.lib().macau_get_link_lambda(m.h, 0)
        print("univariate iter %d: row err %.3e %.3e, mean lambda_beta %.6f vs %.6f" %
              (n_it, e0, e1, lb, ref.lambda_beta[0]))
        assert e0 < TOL_FP32 and e1 < tol_cols
        assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < TOL_FP32
    m.iterate(8)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    print("univariate RMSE gpu %.6f oracle %.6f" % (m.rmse(), ref.rmse))
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


def test_model_univariate_link_and_hyper():
    """beta, mu and the diagonal lambda after two iterations."""
    announce("model univariate link matrix and hyperparameters")
    nrow, ncol, nfeat, K = 200, 80, 30, 6
    (fr, fc, fv), (r, c, y), (rt, ct, yt) = _side_problem(31, nrow, ncol, nfeat, 0.1, False)
    m = GpuModel(2, K, 5.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([rt, ct], 1), yt)
    m.one_prior((fr, fc, fv, nrow, nfeat), lambda_beta=2.0)
    m.prior()
    m.setup(2, 2, SEED, PLAIN, graph=False)
    m.iterate(2)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
    side = [O.csr_pair(fr, fc, fv, nrow, nfeat), None]
    F, Ft = side[0]
    colsq = np.bincount(fc, weights=fv ** 2, minlength=nfeat)
    # replay the oracle step by step to get the prior state after two iterations
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, precision=5.0, side=side, lambda_beta=2.0, burnin=2,
                nsamples=2, seed=SEED, max_iters=2, want_U=True, uni=[True, False])
    ref1 = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, precision=5.0, side=side, lambda_beta=2.0, burnin=2,
                 nsamples=2, seed=SEED, max_iters=1, want_U=True, uni=[True, False])
    # state after iteration 1's prior update, then iteration 2's update from
    # the oracle's own iteration-2 U
    st = O.macau_one_update_prior(ref1.U[0], F, Ft, colsq, np.zeros(K), np.full(K, 10.0), np.zeros((nfeat, K)),
                                  np.zeros((nrow, K)), np.full(K, 2.0), SEED, 0, 0)
    mu, lam, beta, Uhat, lb = O.macau_one_update_prior(ref.U[0], F, Ft, colsq, *st, SEED, 1, 0)
    mu_g, Lam_g = m.hyper(0)
    np.testing.assert_allclose(mu_g, mu, rtol=1e-3, atol=1e-4)
    np.testing.assert_allclose(np.diag(Lam_g), lam, rtol=1e-3)
    assert np.count_nonzero(Lam_g - np.diag(np.diag(Lam_g))) == 0
    np.testing.assert_allclose(m.link(0, nfeat), beta, rtol=1e-3, atol=1e-4)
    assert abs(_lib.lib().macau_get_link_lambda(m.h, 0) - lb.mean()) / lb.mean() < 1e-4


def test_model_univariate_tensor_matches_oracle():
    announce("model univariate tensor")
    rng = np.random.default_rng(41)
    dims, K, nfeat = [150, 40, 8], 8, 20
    Fm = rng.random((dims[0], nfeat)) < 0.15
    fr, fc = np.nonzero(Fm)
    A = Fm.astype(float) @ rng.standard_normal((nfeat, 3)) / 2
    Bm, Cm = rng.standard_normal((dims[1], 3)), rng.standard_normal((dims[2], 3))
    idx = np.stack([rng.integers(0, d, 5000) for d in dims], axis=1)
    val = np.sum(A[idx[:, 0]] * Bm[idx[:, 1]] * Cm[idx[:, 2]], 1) + 0.2 * rng.standard_normal(5000)
    tr, te = idx[:4000], idx[4000:]
    m = GpuModel(3, K, 20.0)
    m.data(tr, val[:4000], dims, te, val[4000:])
    m.one_prior((fr, fc, None, dims[0], nfeat))
    m.prior()
    m.prior()
    m.setup(4, 6, SEED, PLAIN, graph=False)
    modes, mean, nnz = O.tensor_rows(tr, val[:4000], dims)
    ti, tv = O.tensor_test(te, val[4000:], dims)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), dims[0], nfeat), None, None]
    kw = dict(is_tensor=True, precision=20.0, side=side, burnin=4, nsamples=6, seed=SEED, uni=[True, False, False])
    m.iterate(2)
    ref = O.run(modes, mean, nnz, dims, K, ti, tv, max_iters=2, want_U=True, **kw)
    for q in range(3):
        assert row_rel_err(m.latents(q, dims[q]), ref.U[q]) < TOL_FP32
    m.iterate(8)
    ref = O.run(modes, mean, nnz, dims, K, ti, tv, **kw)
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


def test_model_univariate_many_rows():
    """More rows than the sweep keeps in LDS (19456): residuals in global
    memory."""
    announce("model univariate 20000 rows (global-memory sweep)")
    rng = np.random.default_rng(43)
    nrow, ncol, nfeat, K = 20000, 50, 12, 4
    fr = rng.integers(0, nrow, 30000)
    fc = rng.integers(0, nfeat, 30000)
    key = np.unique(fr.astype(np.int64) * nfeat + fc)
    fr, fc = (key // nfeat).astype(np.int32), (key % nfeat).astype(np.int32)
    r = rng.integers(0, nrow, 60000)
    c = rng.integers(0, ncol, 60000)
    key = np.unique(r.astype(np.int64) * ncol + c)
    r, c = (key // ncol).astype(np.int32), (key % ncol).astype(np.int32)
    y = rng.standard_normal(len(r))
    m = GpuModel(2, K, 4.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol])
    m.one_prior((fr, fc, None, nrow, nfeat))
    m.prior()
    m.setup(1, 1, SEED, PLAIN, graph=False)
    m.iterate(2)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, np.zeros((0, 2), np.int32), np.zeros(0), precision=4.0,
                side=side, burnin=1, nsamples=1, seed=SEED, max_iters=2, want_U=True, uni=[True, False])
    assert row_rel_err(m.latents(0, nrow), ref.U[0]) < TOL_FP32
    assert row_rel_err(m.latents(1, ncol), ref.U[1]) < TOL_FP32
    lb = _lib.lib().macau_get_link_lambda(m.h, 0)
    assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < TOL_FP32


def test_model_univariate_long_columns_duplicates():
    """Feature columns with more than 64 rows (the sweep's tail loops, LDS
    form), repeated (row, feature) entries (merged for the sweep; the oracle
    keeps them, as the reference does: At_mul_Bt and add_Acol_mul_bt visit
    both, col_square_sum adds both squares) and empty feature columns."""
    announce("model univariate long columns, repeated entries, empty columns")
    rng = np.random.default_rng(47)
    nrow, ncol, nfeat, K = 2000, 60, 12, 8
    fr = rng.integers(0, nrow, 2400)
    fc = rng.integers(0, nfeat - 3, 2400)   # the last three features stay empty
    fr = np.concatenate([fr, fr[:100]])     # 100 repeated entries
    fc = np.concatenate([fc, fc[:100]])
    fv = rng.standard_normal(len(fr))
    r = rng.integers(0, nrow, 30000)
    c = rng.integers(0, ncol, 30000)
    key = np.unique(r.astype(np.int64) * ncol + c)
    r, c = (key // ncol).astype(np.int32), (key % ncol).astype(np.int32)
    Fd = np.zeros((nrow, nfeat))
    np.add.at(Fd, (fr, fc), fv)
    y = 0.5 * (Fd @ rng.standard_normal(nfeat))[r] + rng.standard_normal(len(r))
    m = GpuModel(2, K, 3.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol])
    m.one_prior((fr, fc, fv, nrow, nfeat))
    m.prior()
    m.setup(1, 2, SEED, PLAIN, graph=False)
    m.iterate(2)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    side = [O.csr_pair(fr, fc, fv, nrow, nfeat), None]
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, np.zeros((0, 2), np.int32), np.zeros(0), precision=3.0,
                side=side, burnin=1, nsamples=2, seed=SEED, max_iters=2, want_U=True, uni=[True, False])
    e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
    lb = _lib.lib().macau_get_link_lambda(m.h, 0)
    print("univariate long/repeated/empty: row err %.3e %.3e, lambda_beta %.6f vs %.6f" %
          (e0, e1, lb, ref.lambda_beta[0]))
    assert e0 < TOL_FP32 and e1 < TOL_FP32
    assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < TOL_FP32
