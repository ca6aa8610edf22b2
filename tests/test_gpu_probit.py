"""GPU parity of the probit noise model (ProbitNoise, noisemodels.h:88-103;
sample_latent_blas_probit, latentprior.cpp:387-414) against the FP64 oracle:
the truncated-normal draw and Phi^{-1} on the device, the augmented model
(per-row latents after one and two iterations, test predictions, AUC) with
the FP32 and the batched split-BF16 samplers, side information, and the
reference's tensor restriction."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from _gpu_helpers import SEED, GpuModel, announce, row_rel_err
from macau_amd import _lib
from macau_amd._lib import check, ptr

pytestmark = pytest.mark.gpu

PLAIN = _lib.SAMPLER_PLAIN_FP32
BATCHED = _lib.SAMPLER_BATCHED
TOL_FP32 = 1e-4
TOL_BATCHED = 1e-4


def test_truncnorm_probe_matches_oracle():
    """Device rand_truncnorm(mean, 1, 0) and Phi^{-1} against the oracle on
    both sides of the inverse-CDF / tail switch (abar = 3)."""
    announce("probit truncnorm probe")
    means = np.concatenate([np.linspace(-10, 5, 1500), [-3.0, -3.0 - 1e-12, -3.0 + 1e-12, 0.0]])
    n = len(means)
    p = np.concatenate([np.logspace(-300, np.log10(0.5), n // 2), 1 - np.logspace(-16, np.log10(0.5), n - n // 2)])
    out, inv = np.zeros(n), np.zeros(n)
    tg = O.tag(O.TAG_PROBIT, 1)
    check(_lib.lib().macau_truncnorm_probe(SEED, 100, 3, tg, n, ptr(means), ptr(p), ptr(out), ptr(inv)))
    ref = np.array([O.truncnorm0(m, SEED, 100 + j, 3, tg) for j, m in enumerate(means)])
    ref_inv = np.array([O.inv_norm_cdf(x) for x in p])
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(inv, ref_inv, rtol=1e-13, atol=1e-15)
    assert np.all(out >= 0.0)


def _probit_problem(seed, nrow, ncol, rank, density):
    rng = np.random.default_rng(seed)
    A, B = rng.standard_normal((nrow, rank)), rng.standard_normal((ncol, rank))
    r, c = np.nonzero(rng.random((nrow, ncol)) < density)
    y = (np.sum(A[r] * B[c], 1) > 0).astype(float)
    perm = rng.permutation(len(r))
    nte = len(r) // 5
    te, tr = perm[:nte], perm[nte:]
    return (r[tr], c[tr], y[tr]), (r[te], c[te], y[te])


@pytest.mark.parametrize("K,sampler", [(8, "plain"), (32, "batched"), (48, "batched")])
def test_model_probit_matches_oracle(K, sampler):
    announce("model probit K=%d sampler=%s" % (K, sampler))
    nrow, ncol = 400, 250
    (r, c, y), (rt, ct, yt) = _probit_problem(11, nrow, ncol, 3, 0.15)
    kind, tol = (PLAIN, TOL_FP32) if sampler == "plain" else (BATCHED, TOL_BATCHED)
    m = GpuModel(2, K, "probit")
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([rt, ct], 1), yt)
    m.prior()
    m.prior()
    m.setup(4, 8, SEED, kind)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
    kw = dict(probit=True, burnin=4, nsamples=8, seed=SEED)
    for n_it in (1, 2):
        m.iterate(1)
        ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=n_it, want_U=True, **kw)
        e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
        print("probit iter %d: row err %.3e %.3e" % (n_it, e0, e1))
        assert e0 < tol and e1 < tol
        # before burn-in ends the stored prediction is the one-sample Phi(u.v)
        np.testing.assert_allclose(m.predictions(len(tv)), ref.pred, rtol=10 * tol, atol=10 * tol)
        a1, aa = m.auc()
        assert abs(a1 - ref.auc1) < 0.01 and abs(aa - ref.auc) < 0.01
    m.iterate(10)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    a1, aa = m.auc()
    print("probit AUC gpu %.5f oracle %.5f" % (aa, ref.auc))
    assert aa > 0.85
    assert abs(aa - ref.auc) < 0.02
    assert math.isnan(m.rmse()) or m.rmse() >= 0.0


def test_model_probit_side_matches_oracle():
    """Probit with side information on the rows (latentprior.cpp:296-303:
    the sampler's prior mean is mu + Uhat_n), F^T F branch."""
    announce("model probit side information")
    rng = np.random.default_rng(13)
    nrow, ncol, K, nfeat = 300, 100, 8, 40
    Fm = rng.random((nrow, nfeat)) < 0.1
    fr, fc = np.nonzero(Fm)
    W = rng.standard_normal((nfeat, 3))
    A = Fm.astype(float) @ W / 2
    B = rng.standard_normal((ncol, 3))
    r, c = np.nonzero(rng.random((nrow, ncol)) < 0.2)
    y = (np.sum(A[r] * B[c], 1) > 0).astype(float)
    perm = rng.permutation(len(r))
    te, tr = perm[:len(r) // 5], perm[len(r) // 5:]
    m = GpuModel(2, K, "probit")
    m.data(np.stack([r[tr], c[tr]], 1), y[tr], [nrow, ncol], np.stack([r[te], c[te]], 1), y[te])
    m.prior(side=(fr, fc, None, nrow, nfeat), lambda_beta=5.0, tol=1e-6, max_ff_size=10000)
    m.prior()
    m.setup(3, 5, SEED, PLAIN, graph=False)
    modes, mean, nnz = O.matrix_rows(r[tr], c[tr], y[tr], nrow, ncol)
    ti, tv = O.matrix_test(r[te], c[te], y[te], nrow, ncol)
    side = [O.csr_pair(fr, fc, np.ones(len(fr)), nrow, nfeat), None]
    kw = dict(probit=True, side=side, lambda_beta=5.0, tol=1e-6, burnin=3, nsamples=5, seed=SEED,
              use_ftf=[True, False])
    m.iterate(2)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, max_iters=2, want_U=True, **kw)
    e0, e1 = row_rel_err(m.latents(0, nrow), ref.U[0]), row_rel_err(m.latents(1, ncol), ref.U[1])
    print("probit side iter 2: row err %.3e %.3e" % (e0, e1))
    assert e0 < TOL_FP32 and e1 < TOL_FP32
    m.iterate(6)
    ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, **kw)
    _, aa = m.auc()
    assert abs(aa - ref.auc) < 0.02


def test_probit_tensor_rejected():
    """BPMFPrior::sample_latents(ProbitNoise&, TensorData&) throws
    (latentprior.cpp:112-116)."""
    announce("probit tensor rejected")
    L = _lib.lib()
    h = L.macau_make_probit(3, 4)
    assert h
    try:
        idx = np.asfortranarray(np.array([[0, 0, 0], [1, 1, 1], [2, 0, 1]], dtype=np.int32))
        dims = np.array([3, 2, 2], dtype=np.int32)
        check(L.macau_set_relation_data(h, ptr(idx), 3, ptr(np.array([1.0, 0.0, 1.0])), 3, ptr(dims)))
        for _ in range(3):
            check(L.macau_add_bpmf_prior(h))
        check(L.macau_set_verbose(h, 0))
        assert L.macau_init(h) != 0
        assert "Unimplemented: sample_latents" in _lib.last_error()
    finally:
        L.macau_free(h)


def test_auc_only_for_probit():
    announce("auc getter on a gaussian model")
    m = GpuModel(2, 4, 2.0)
    a = ctypes.c_double()
    assert _lib.lib().macau_get_auc(m.h, ctypes.byref(a), ctypes.byref(a)) != 0
    assert "probit" in _lib.last_error()
