"""Parity of the HIP path with the FP64 oracle on identical inputs and draws.

Tolerances: every sampled latent row, link matrix and
hyperparameter of the FP32 path within 1e-4 relative of the FP64 oracle
(row-wise: ||u_gpu - u_ref|| / ||u_ref||); FP64 kernels within 1e-9;
integer RNG words bit-exact; posterior-mean test RMSE within 0.5%.
"""
import numpy as np
import pytest

import oracle as O
from _gpu_helpers import (SEED, GpuModel, announce, f32, gpu_blockcg, gpu_cond_normal_wishart,
                          gpu_sample_latents, power_law_rows, rel_err, row_rel_err, spd)
from macau_amd import _lib
from macau_amd._lib import ptr, check

pytestmark = pytest.mark.gpu

PLAIN = _lib.SAMPLER_PLAIN_FP32
TOL_FP32 = 1e-4


@pytest.mark.parametrize("tag", [O.tag(O.TAG_LATENT, 0), O.tag(O.TAG_WISHART_Z, 1), O.tag(O.TAG_BETA_F, 3)])
def test_rng_words_and_normals(tag):
    announce("rng probe tag=%x" % tag)
    n = 1024
    for c0, it in [(0, 0), (12345, 7), (2**31 + 5, 99)]:
        bits = np.zeros(n, dtype=np.uint32)
        z32 = np.zeros(n, dtype=np.float32)
        z64 = np.zeros(n)
        check(_lib.lib().macau_rng_probe(SEED, c0, it, tag, n, ptr(bits), ptr(z32), ptr(z64)))
        ref_bits = np.concatenate([O.philox([c0, b, it, tag], [SEED & 0xFFFFFFFF, SEED >> 32]) for b in range(n // 4)])
        np.testing.assert_array_equal(bits, ref_bits)
        ref = O.normals(SEED, c0, n, it, tag)
        np.testing.assert_allclose(z64, ref, rtol=0, atol=1e-13)
        np.testing.assert_allclose(z32, ref, rtol=0, atol=2e-5)


def test_gamma_draws():
    announce("gamma probe")
    shapes = np.array([0.2, 0.5, 1.0, 1.5, 3.0, 16.0, 250.0, 5e4, 9e6] * 8)
    out = np.zeros(len(shapes))
    tg = O.tag(O.TAG_WISHART_CHI, 2)
    check(_lib.lib().macau_gamma_probe(SEED, ptr(shapes), len(shapes), 5, tg, ptr(out)))
    ref = np.array([O.gamma(SEED, s, j, 5, tg) for j, s in enumerate(shapes)])
    np.testing.assert_allclose(out, ref, rtol=1e-12)


def _matrix_case(seed, K, nrows, ncols, nnz, heavy=()):
    rng = np.random.default_rng(seed)
    r, c, y = power_law_rows(rng, nrows, ncols, nnz, heavy)
    modes, mean, _ = O.matrix_rows(r, c, y, nrows, ncols)
    V = f32(rng.standard_normal((ncols, K)) * 0.5)
    Lam = spd(rng, K, 3.0)
    mu = rng.standard_normal(K) * 0.3
    return rng, modes, mean, V, Lam, mu


@pytest.mark.parametrize("K", [2, 4, 10, 16, 32, 64, 128])
def test_plain_sampler_matrix(K):
    announce("plain sampler matrix K=%d" % K)
    rng, modes, mean, V, Lam, mu = _matrix_case(K, K, 300, 2000, 6000, heavy=(1500, 0, 3))
    for mode_rows, mode in [(modes[0], 0)]:
        got = gpu_sample_latents(K, mode_rows, mean, [V], 2.5, mu, Lam, SEED, 3, mode, PLAIN)
        ref = O.sample_latents(K, mode_rows, mean, [V], 2.5, mu, Lam, SEED, 3, mode)
        assert row_rel_err(got, ref) < TOL_FP32


def test_plain_sampler_item_side():
    announce("plain sampler item side")
    K = 32
    rng, modes, mean, V, Lam, mu = _matrix_case(1, K, 3000, 200, 30000)
    U = f32(rng.standard_normal((3000, K)) * 0.5)
    got = gpu_sample_latents(K, modes[1], mean, [U], 1.3, mu, Lam, SEED, 11, 1, PLAIN)
    ref = O.sample_latents(K, modes[1], mean, [U], 1.3, mu, Lam, SEED, 11, 1)
    assert row_rel_err(got, ref) < TOL_FP32


@pytest.mark.parametrize("K", [8, 16])
def test_plain_sampler_tensor(K):
    announce("plain sampler tensor K=%d" % K)
    rng = np.random.default_rng(K + 100)
    dims = [300, 80, 12]
    idx = np.stack([rng.integers(0, d, 5000) for d in dims], axis=1)
    vals = rng.standard_normal(5000)
    modes, mean, _ = O.tensor_rows(idx, vals, dims)
    fs = [f32(rng.standard_normal((d, K)) * 0.7) for d in dims]
    Lam, mu = spd(rng, K, 2.0), rng.standard_normal(K) * 0.2
    for mode in range(3):
        others = [fs[q] for q in range(3) if q != mode]
        got = gpu_sample_latents(K, modes[mode], mean, others, 4.0, mu, Lam, SEED, 1, mode, PLAIN)
        ref = O.sample_latents(K, modes[mode], mean, others, 4.0, mu, Lam, SEED, 1, mode)
        assert row_rel_err(got, ref) < TOL_FP32


def test_plain_sampler_per_row_mean():
    announce("plain sampler per-row mean")
    K = 32
    rng, modes, mean, V, Lam, mu = _matrix_case(7, K, 400, 300, 4000)
    mu_rows = f32(mu + rng.standard_normal((400, K)) * 0.2)
    got = gpu_sample_latents(K, modes[0], mean, [V], 2.0, mu, Lam, SEED, 4, 0, PLAIN, mu_rows=mu_rows)
    ref = O.sample_latents(K, modes[0], mean, [V], 2.0, mu, Lam, SEED, 4, 0, mu_rows=mu_rows)
    assert row_rel_err(got, ref) < TOL_FP32


@pytest.mark.parametrize("K,N", [(3, 50), (10, 7000), (32, 20000), (64, 3000), (65, 800), (100, 500), (128, 4000)])
def test_normal_wishart(K, N):
    announce("normal-wishart K=%d N=%d" % (K, N))
    rng = np.random.default_rng(K)
    U = rng.standard_normal((N, K)) * 0.5 + 0.1
    T = spd(rng, K, 0.7)
    mu0 = np.zeros(K)
    for it in (0, 9):
        mu, Lam = gpu_cond_normal_wishart(U, mu0, 2.0, T, K + 3, SEED, it, 1)
        mur, Lamr = O.cond_normal_wishart(U, mu0, 2.0, T, K + 3, SEED, it, 1)
        assert rel_err(Lam, Lamr) < 1e-9
        assert rel_err(mu, mur) < 1e-9


def _feat(kat):
    f = kat["feat"]
    return f["rows"], f["cols"], f["nrow"], f["ncol"]


def test_blockcg_reference_known_answers(kat):
    announce("blockcg known answers")
    r, c, nr, nc = _feat(kat)
    c1 = kat["blockcg_1"]
    X, it = gpu_blockcg(r, c, None, nr, nc, c1["reg"], np.array(c1["B"]), c1["tol"])
    np.testing.assert_allclose(X, np.array(c1["X"]), rtol=1e-6)
    assert it <= c1["max_iter"]
    c3 = kat["blockcg_3"]
    X, it = gpu_blockcg(r, c, None, nr, nc, c3["reg"], np.array(c3["B"]), c3["tol"])
    np.testing.assert_allclose(X, np.array(c3["X"]), rtol=1e-6)


@pytest.mark.parametrize("nrhs,binary", [(1, True), (5, False), (8, False), (16, True), (20, False), (32, True), (32, False), (40, True)])
def test_blockcg_vs_oracle(nrhs, binary):
    announce("blockcg nrhs=%d" % nrhs)
    rng = np.random.default_rng(nrhs)
    nrow, ncol, nnz = 3000, 2000, 30000
    r = rng.integers(0, nrow, nnz)
    c = np.minimum((rng.pareto(1.5, nnz) * 30).astype(int), ncol - 1)
    key = np.unique(r.astype(np.int64) * ncol + c)
    r, c = (key // ncol).astype(np.int32), (key % ncol).astype(np.int32)
    v = None if binary else rng.random(len(r)) + 0.5
    B = rng.standard_normal((nrhs, ncol))
    X, it = gpu_blockcg(r, c, v, nrow, ncol, 5.0, B, 1e-6)
    F, Ft = O.csr_pair(r, c, np.ones(len(r)) if v is None else v, nrow, ncol)
    Xr, itr = O.blockcg(F, Ft, 5.0, B.T, 1e-6)
    assert rel_err(X, Xr.T) < 1e-6
    # the stopping test compares a residual norm against tol; FP64 sums in a
    # different order can move that crossing by an iteration or two
    assert abs(it - itr) <= 2


@pytest.mark.parametrize("ones", [False, True])
def test_blockcg_dense_block_vs_oracle(ones):
    """32 right-hand sides over a binary F with 40 dense columns (a third of
    the rows each) and 0-20 sparse nonzeros per row: the dense-block SpMM
    products, with rows whose remaining nonzeros take one, two or three
    batches; nrow is not a multiple of the 32-row tiles or the 256-row
    chunks.  ones: the same F given as values 1.0."""
    announce("blockcg dense block ones=%s" % ones)
    rng = np.random.default_rng(77)
    nrow, ncol, nd = 1500, 3000, 40
    dense = rng.choice(ncol, nd, replace=False)
    rows, cols = [], []
    for r in range(nrow):
        d = dense[rng.random(nd) < 0.33]
        k = rng.integers(0, 21)
        sp = rng.choice(np.setdiff1d(np.arange(ncol), dense)[:1200], k, replace=False)
        for c in np.concatenate([d, sp]):
            rows.append(r)
            cols.append(int(c))
    r = np.array(rows, dtype=np.int32)
    c = np.array(cols, dtype=np.int32)
    v = np.ones(len(r)) if ones else None
    B = rng.standard_normal((32, ncol))
    X, it = gpu_blockcg(r, c, v, nrow, ncol, 5.0, B, 1e-6)
    F, Ft = O.csr_pair(r, c, np.ones(len(r)), nrow, ncol)
    Xr, itr = O.blockcg(F, Ft, 5.0, B.T, 1e-6)
    assert rel_err(X, Xr.T) < 1e-6
    assert abs(it - itr) <= 2


# ------------------------------------------------------------ model level
def _bpmf_problem(seed, nrow, ncol, K, density, noise):
    rng = np.random.default_rng(seed)
    U, V = rng.standard_normal((nrow, K)), rng.standard_normal((ncol, K))
    mask = rng.random((nrow, ncol)) < density
    r, c = np.nonzero(mask)
    y = np.sum(U[r] * V[c], 1) + noise * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    nte = len(r) // 5
    return (r[perm[nte:]], c[perm[nte:]], y[perm[nte:]]), (r[perm[:nte]], c[perm[:nte]], y[perm[:nte]])


@pytest.mark.parametrize("precision", [5.0, "adaptive"])
def test_model_iterations_match_oracle(precision):
    announce("model bpmf precision=%s" % precision)
    nrow, ncol, K = 500, 300, 8
    (r, c, y), (rt, ct, yt) = _bpmf_problem(3, nrow, ncol, 6, 0.1, 0.5)
    m = GpuModel(2, K, precision)
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([rt, ct], 1), yt)
    m.prior()
    m.prior()
    m.setup(5, 10, SEED, PLAIN)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
    adaptive = precision == "adaptive"
    kw = dict(adaptive=adaptive, precision=1.0 if adaptive else precision, burnin=5, nsamples=10, seed=SEED)
    for n_it in (1, 2):
        m.iterate(1)
        ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=n_it, want_U=True, **kw)
        assert row_rel_err(m.latents(0, nrow), ref.U[0]) < TOL_FP32
        assert row_rel_err(m.latents(1, ncol), ref.U[1]) < TOL_FP32
        assert abs(m.alpha() - ref.alpha) / ref.alpha < TOL_FP32
    m.iterate(13)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


def test_model_tensor_matches_oracle():
    announce("model tensor")
    rng = np.random.default_rng(5)
    dims, K = [200, 60, 10], 8
    A = [rng.standard_normal((d, 3)) for d in dims]
    idx = np.stack([rng.integers(0, d, 6000) for d in dims], axis=1)
    val = np.sum(A[0][idx[:, 0]] * A[1][idx[:, 1]] * A[2][idx[:, 2]], 1) + 0.3 * rng.standard_normal(6000)
    tr, te = idx[:5000], idx[5000:]
    m = GpuModel(3, K, "adaptive")
    m.data(tr, val[:5000], dims, te, val[5000:])
    for _ in range(3):
        m.prior()
    m.setup(5, 10, SEED, PLAIN)
    modes, mean, nnz = O.tensor_rows(tr, val[:5000], dims)
    ti, tv = O.tensor_test(te, val[5000:], dims)
    kw = dict(is_tensor=True, adaptive=True, burnin=5, nsamples=10, seed=SEED)
    m.iterate(1)
    ref = O.run(modes, mean, nnz, dims, K, ti, tv, max_iters=1, want_U=True, **kw)
    for q in range(3):
        assert row_rel_err(m.latents(q, dims[q]), ref.U[q]) < TOL_FP32
    assert abs(m.alpha() - ref.alpha) / ref.alpha < TOL_FP32
    m.iterate(14)
    ref = O.run(modes, mean, nnz, dims, K, ti, tv, **kw)
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


def test_model_macau_matches_oracle():
    announce("model macau side information")
    rng = np.random.default_rng(9)
    nrow, ncol, K, nfeat = 400, 120, 8, 500
    Fd = rng.random((nrow, nfeat)) < 0.02
    fr, fc = np.nonzero(Fd)
    W = rng.standard_normal((nfeat, 3))
    Ut = Fd.astype(float) @ W / 2
    Vt = rng.standard_normal((ncol, 3))
    mask = rng.random((nrow, ncol)) < 0.15
    r, c = np.nonzero(mask)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    m = GpuModel(2, K, 10.0)
    m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
    m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6)
    m.prior()
    m.setup(4, 6, SEED, PLAIN, graph=False)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    kw = dict(precision=10.0, side=side, lambda_beta=5.0, tol=1e-6, burnin=4, nsamples=6, seed=SEED)
    m.iterate(1)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=1, want_U=True, **kw)
    assert row_rel_err(m.latents(0, nrow), ref.U[0]) < TOL_FP32
    assert row_rel_err(m.latents(1, ncol), ref.U[1]) < TOL_FP32
    lb = _lib.lib().macau_get_link_lambda(m.h, 0)
    assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < TOL_FP32
    m.iterate(9)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


@pytest.mark.parametrize("graph", [False, True])
def test_model_macau_dense_block_matches_oracle(graph):
    """K = 32 with binary side information that has 40 dense columns, sparse
    columns among the first 1000 and 2000 featureless ones: the block CG's
    dense-block products and its K P = lambda P shortcut inside the Gibbs
    loop (eager, and as replayed chunks of CG iterations)."""
    announce("model macau dense block graph=%s" % graph)
    rng = np.random.default_rng(21)
    nrow, ncol, K, nfeat = 1200, 150, 32, 3000
    Fd = np.zeros((nrow, nfeat), dtype=bool)
    dense = rng.choice(1000, 40, replace=False)
    Fd[:, dense] = rng.random((nrow, 40)) < 0.3
    Fd[:, :1000] |= rng.random((nrow, 1000)) < 0.004
    fr, fc = np.nonzero(Fd)
    W = rng.standard_normal((nfeat, 3)) * 0.3
    Ut = Fd.astype(float) @ W
    Vt = rng.standard_normal((ncol, 3))
    mask = rng.random((nrow, ncol)) < 0.1
    r, c = np.nonzero(mask)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    m = GpuModel(2, K, 10.0)
    m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
    m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6)
    m.prior()
    m.setup(4, 6, SEED, PLAIN, graph=graph)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    kw = dict(precision=10.0, side=side, lambda_beta=5.0, tol=1e-6, burnin=4, nsamples=6, seed=SEED)
    m.iterate(1)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=1, want_U=True, **kw)
    assert row_rel_err(m.latents(0, nrow), ref.U[0]) < TOL_FP32
    assert row_rel_err(m.latents(1, ncol), ref.U[1]) < TOL_FP32
    lb = _lib.lib().macau_get_link_lambda(m.h, 0)
    assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < TOL_FP32
    m.iterate(4)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=5, **kw)
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


@pytest.mark.parametrize("solver", ["spectral", "cholesky"])
@pytest.mark.parametrize("nfeat,binary", [(70, True), (130, False), (33, False), (700, True)])
def test_model_macau_ftf_matches_oracle(nfeat, binary, solver, monkeypatch):
    """The F^T F branch of sample_beta (latentprior.cpp:280-289), F^T F formed
    on the GPU, against the oracle's Cholesky solve.  "cholesky": blocked
    Cholesky of F^T F + lambda_beta I and two blocked solves every iteration;
    "spectral": one eigendecomposition of F^T F at setup and two products per
    iteration (700 features: six inner-dimension chunks).  Feature counts
    leave partial 32-column blocks and 128-row tiles."""
    announce("model macau F^T F branch nfeat=%d binary=%d solver=%s" % (nfeat, binary, solver))
    monkeypatch.setenv("MACAU_AMD_FTF_SOLVER", solver)
    rng = np.random.default_rng(nfeat)
    nrow, ncol, K = 300, 90, 8
    Fm = rng.random((nrow, nfeat)) < 0.08
    fr, fc = np.nonzero(Fm)
    fv = np.ones(len(fr)) if binary else rng.standard_normal(len(fr))
    Ut = rng.standard_normal((nrow, 3))
    Vt = rng.standard_normal((ncol, 3))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.2)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    m = GpuModel(2, K, 10.0)
    m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
    m.prior(side=(fr, fc, None if binary else fv, nrow, nfeat), lambda_beta=5.0, tol=1e-6, max_ff_size=10000)
    m.prior()
    m.setup(4, 6, SEED, PLAIN, graph=False)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = [O.csr_pair(fr, fc, fv, nrow, nfeat), None]
    kw = dict(precision=10.0, side=side, lambda_beta=5.0, tol=1e-6, burnin=4, nsamples=6, seed=SEED,
              use_ftf=[True, False])
    m.iterate(2)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=2, want_U=True, **kw)
    e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
    lb = _lib.lib().macau_get_link_lambda(m.h, 0)
    print("F^T F model iter 2: row err %.3e %.3e, lambda_beta %.6f vs %.6f" % (e0, e1, lb, ref.lambda_beta[0]))
    assert e0 < TOL_FP32 and e1 < TOL_FP32
    assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < TOL_FP32
    assert _lib.lib().macau_get_last_cg_iters(m.h, 0) == 0
    assert _lib.lib().macau_get_ftf_solver(m.h, 0) == (2 if solver == "spectral" else 1)
    assert _lib.lib().macau_get_ftf_solver(m.h, 1) == 0
    m.iterate(8)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005


@pytest.mark.parametrize("solver", ["spectral", "cholesky"])
def test_model_macau_ftf_not_positive_definite(solver, monkeypatch):
    """F^T F + lambda_beta I not positive definite (a negative initial
    lambda_beta): the F^T F branch reports chol_decomp's failure (chol.cpp:21-28)
    with either solver -- a non-positive pivot, or a non-positive
    d_j + lambda_beta -- after its kernels have finished normally."""
    announce("model macau F^T F not positive definite solver=%s" % solver)
    monkeypatch.setenv("MACAU_AMD_FTF_SOLVER", solver)
    rng = np.random.default_rng(5)
    nrow, ncol, K, nfeat = 120, 60, 8, 40
    fr, fc = np.nonzero(rng.random((nrow, nfeat)) < 0.1)
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.2)
    y = rng.standard_normal(len(r))
    m = GpuModel(2, K, 10.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([r[:10], c[:10]], 1), y[:10])
    m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=-1e6, tol=1e-6, max_ff_size=10000)
    m.prior()
    m.setup(2, 2, SEED, PLAIN, graph=False)
    with pytest.raises(RuntimeError, match="Cholesky decomp failed"):
        m.iterate(1)


@pytest.mark.parametrize("ftf", [True, False])
def test_model_macau_wide_latent(ftf):
    """K = 70 (more than 64 right-hand sides in the SpMMs of sample_beta and
    compute_uhat), with the F^T F branch or the block CG (blocks 32 + 38)."""
    announce("model macau K=70 ftf=%d" % ftf)
    rng = np.random.default_rng(70 + ftf)
    nrow, ncol, K, nfeat = 200, 80, 70, 40
    Fm = rng.random((nrow, nfeat)) < 0.1
    fr, fc = np.nonzero(Fm)
    Ut = Fm.astype(float) @ rng.standard_normal((nfeat, 3)) / 2
    Vt = rng.standard_normal((ncol, 3))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.2)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    m = GpuModel(2, K, 10.0)
    m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
    m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6, max_ff_size=10000 if ftf else None)
    m.prior()
    m.setup(3, 4, SEED, PLAIN, graph=False)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    kw = dict(precision=10.0, side=side, lambda_beta=5.0, tol=1e-6, burnin=3, nsamples=4, seed=SEED,
              use_ftf=[ftf, False])
    m.iterate(2)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=2, want_U=True, **kw)
    e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
    print("macau K=70 ftf=%d iter 2: row err %.3e %.3e" % (ftf, e0, e1))
    # the block CG stops at tolerance 1e-6: its solution error bounds the parity
    tol = TOL_FP32 if ftf else 1e-3
    assert e0 < tol and e1 < tol
    lb = _lib.lib().macau_get_link_lambda(m.h, 0)
    assert abs(lb - ref.lambda_beta[0]) / ref.lambda_beta[0] < tol


@pytest.mark.parametrize("kind,K", [("ftf", 8), ("univariate", 8), ("cg", 8), ("cg", 70)])
def test_macau_graph_replay_matches_eager(kind, K):
    """A Macau chain with graph replay on, against the same chain launched
    eagerly, with host reads of the latents between iterations: the F^T F
    branch (spectral solve) and the univariate prior are captured with the
    whole iteration; the block CG replays its 8-iteration chunks as graphs
    between its host checks of the convergence flags (K = 70: two CG blocks,
    32 and 38 right-hand sides, each with its own graph)."""
    announce("macau graph replay kind=%s K=%d" % (kind, K))
    rng = np.random.default_rng(91)
    nrow, ncol, nfeat = 300, 90, 130
    Fm = rng.random((nrow, nfeat)) < 0.08
    fr, fc = np.nonzero(Fm)
    Ut = rng.standard_normal((nrow, 3))
    Vt = rng.standard_normal((ncol, 3))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.2)
    y = np.sum(Ut[r] * Vt[c], 1) + 0.3 * rng.standard_normal(len(r))
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    models = []
    for graph in (True, False):
        m = GpuModel(2, K, 10.0)
        m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
        if kind == "ftf":
            m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6, max_ff_size=10000)
        elif kind == "cg":
            m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6)
        else:
            m.one_prior((fr, fc, None, nrow, nfeat), lambda_beta=5.0)
        m.prior()
        m.setup(3, 4, SEED, PLAIN, graph=graph)
        models.append(m)
    g, e = models
    for _ in range(7):
        g.iterate(1)
        e.iterate(1)
        for mode, n in ((0, nrow), (1, ncol)):
            np.testing.assert_allclose(g.latents(mode, n), e.latents(mode, n), rtol=1e-6, atol=1e-6)
        if kind == "cg":
            L = _lib.lib()
            assert L.macau_get_last_cg_iters(g.h, 0) == L.macau_get_last_cg_iters(e.h, 0) > 0
    assert abs(g.rmse() - e.rmse()) <= 1e-6 * abs(e.rmse())


# ------------------------------------------------------------ batched sampler
BATCHED = _lib.SAMPLER_BATCHED
# Split-BF16 Gram (hi*hi + hi*lo + lo*hi, FP32 accumulate): the row error
# against the FP64 oracle is printed and held to the FP32 path's 1e-4 (the
# largest measured here is 2e-5); the posterior-mean RMSE gate (0.5%) is
# checked at model level.
TOL_BATCHED = 1e-4


@pytest.mark.parametrize("K", [4, 10, 16, 17, 24, 32, 33, 48, 64, 65, 100, 128])
def test_batched_sampler_matrix(K):
    announce("batched sampler matrix K=%d" % K)
    # heavy rows above the 2048-nnz segment length, rows at the boundary, an
    # empty row, odd row count
    rng, modes, mean, V, Lam, mu = _matrix_case(K + 7, K, 301, 5000, 9000, heavy=(4000, 2049, 2048, 0, 3000))
    got = gpu_sample_latents(K, modes[0], mean, [V], 2.5, mu, Lam, SEED, 3, 0, BATCHED)
    ref = O.sample_latents(K, modes[0], mean, [V], 2.5, mu, Lam, SEED, 3, 0)
    err = row_rel_err(got, ref)
    print("batched K=%d max row rel err %.3e" % (K, err))
    assert err < TOL_BATCHED


@pytest.mark.parametrize("K", [32, 64, 128])
def test_batched_sampler_item_side_and_tensor(K):
    announce("batched sampler item side / tensor K=%d" % K)
    rng, modes, mean, V, Lam, mu = _matrix_case(2, K, 3000, 200, 60000)
    U = f32(rng.standard_normal((3000, K)) * 0.5)
    got = gpu_sample_latents(K, modes[1], mean, [U], 1.3, mu, Lam, SEED, 11, 1, BATCHED)
    ref = O.sample_latents(K, modes[1], mean, [U], 1.3, mu, Lam, SEED, 11, 1)
    assert row_rel_err(got, ref) < TOL_BATCHED
    dims = [400, 90, 12]
    idx = np.stack([rng.integers(0, d, 8000) for d in dims], axis=1)
    vals = rng.standard_normal(8000)
    tmodes, tmean, _ = O.tensor_rows(idx, vals, dims)
    fs = [f32(rng.standard_normal((d, K)) * 0.7) for d in dims]
    for mode in range(3):
        others = [fs[q] for q in range(3) if q != mode]
        got = gpu_sample_latents(K, tmodes[mode], tmean, others, 4.0, mu, Lam, SEED, 1, mode, BATCHED)
        ref = O.sample_latents(K, tmodes[mode], tmean, others, 4.0, mu, Lam, SEED, 1, mode)
        assert row_rel_err(got, ref) < TOL_BATCHED


@pytest.mark.parametrize("K", [32, 64, 128])
def test_batched_sampler_per_row_mean(K):
    announce("batched sampler per-row mean K=%d" % K)
    rng, modes, mean, V, Lam, mu = _matrix_case(8, K, 400, 3000, 8000, heavy=(2500,))
    mu_rows = f32(mu + rng.standard_normal((400, K)) * 0.2)
    got = gpu_sample_latents(K, modes[0], mean, [V], 2.0, mu, Lam, SEED, 4, 0, BATCHED, mu_rows=mu_rows)
    ref = O.sample_latents(K, modes[0], mean, [V], 2.0, mu, Lam, SEED, 4, 0, mu_rows=mu_rows)
    assert row_rel_err(got, ref) < TOL_BATCHED


@pytest.mark.parametrize("precision,K", [(5.0, 32), ("adaptive", 32), ("adaptive", 64), ("adaptive", 10),
                                         ("adaptive", 128)])
def test_batched_model_matches_oracle(precision, K):
    announce("batched model precision=%s K=%d" % (precision, K))
    nrow, ncol = 3000, 3000
    rng = np.random.default_rng(21)
    r, c, y = power_law_rows(rng, nrow, ncol, 90000, heavy=(2600, 2300))  # two heavy rows of two segments
    y = np.round(y * 2) / 2
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 10], perm[len(r) // 10:]
    m = GpuModel(2, K, precision)
    m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
    m.prior()
    m.prior()
    m.setup(10, 20, SEED, BATCHED)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    adaptive = precision == "adaptive"
    kw = dict(adaptive=adaptive, precision=1.0 if adaptive else precision, burnin=10, nsamples=20, seed=SEED)
    m.iterate(1)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=1, want_U=True, **kw)
    e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
    ea = abs(m.alpha() - ref.alpha) / ref.alpha
    print("batched model iter 1: row err %.3e %.3e, alpha err %.3e" % (e0, e1, ea))
    assert e0 < TOL_BATCHED and e1 < TOL_BATCHED
    assert ea < 2e-5
    m.iterate(29)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    print("batched model RMSE %.6f vs oracle %.6f" % (m.rmse(), ref.rmse))
    assert abs(m.rmse() - ref.rmse) / ref.rmse < 0.005
