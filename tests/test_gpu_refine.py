"""One step of iterative refinement for ill-conditioned rows of the batched
sampler (gated by the probe condition estimate of the Jacobi-scaled
precision matrix).

Rows of 2000 to 27000 nonzeros whose factor columns are nearly collinear
(one direction of the other mode's factors scaled by 0.03, condition numbers
~1000) at noise precision 10: without refinement the split-BF16 sampler's
rows differ from the FP64 oracle by more than 1e-4; with it, every row with
at least max(K, 128) nonzeros is within 1e-4.  Rows with fewer nonzeros
are never probed or queued and keep their values.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from _gpu_helpers import SEED, GpuModel, announce, f32, gpu_sample_latents, row_rel_err, spd
from macau_amd import _lib
from macau_amd._lib import check

pytestmark = pytest.mark.gpu


def _collinear_case(K, seed):
    rng = np.random.default_rng(seed)
    ncol = 30000
    Q = np.linalg.qr(rng.standard_normal((K, K)))[0]
    s = np.ones(K)
    s[0] = 0.03
    V = f32((rng.standard_normal((ncol, K)) * s) @ Q.T * 0.5)
    degs = [2000, 4000, 8000, 16000, 27000, 27000, 20, 45, 300, 100]
    r = np.concatenate([np.full(d, i) for i, d in enumerate(degs)])
    c = np.concatenate([rng.choice(ncol, d, replace=False) for d in degs])
    u_true = rng.standard_normal((len(degs), K))
    y = np.sum(u_true[r] * V[c], 1) + 0.3 * rng.standard_normal(len(r))
    modes, mean, _ = O.matrix_rows(r, c, y, len(degs), ncol)
    return modes[0], mean, V, spd(rng, K, 1.0), np.zeros(K), np.array(degs)


@pytest.mark.parametrize("K", [32, 64])
def test_refinement_brings_ill_conditioned_heavy_rows_within_1e4(K):
    announce("refinement of ill-conditioned heavy rows K=%d" % K)
    rows, mean, V, Lam, mu, degs = _collinear_case(K, K + 1)
    alpha = 10.0
    ref = O.sample_latents(K, rows, mean, [V], alpha, mu, Lam, SEED, 3, 0)
    err = {}
    for kind in (_lib.SAMPLER_BATCHED_NOREFINE, _lib.SAMPLER_BATCHED):
        got = gpu_sample_latents(K, rows, mean, [V], alpha, mu, Lam, SEED, 3, 0, kind)
        err[kind] = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        print("K=%d %s: row errors %s" % (K, "refined" if kind == _lib.SAMPLER_BATCHED else "unrefined",
                                          " ".join("%d:%.1e" % (d, e) for d, e in zip(degs, err[kind]))))
    heavy = degs >= 2000
    assert err[_lib.SAMPLER_BATCHED_NOREFINE][heavy].max() > 1e-4
    queued = degs >= max(K, 128)
    assert err[_lib.SAMPLER_BATCHED][queued].max() <= 1e-4
    np.testing.assert_array_equal(err[_lib.SAMPLER_BATCHED][~queued], err[_lib.SAMPLER_BATCHED_NOREFINE][~queued])


@pytest.mark.parametrize("K,lo,hi", [(32, 0.1, 1.0), (64, 0.2, 1.5)])
def test_spread_component_scales_are_not_queued(K, lo, hi):
    """Rows whose latent components have very different scales but no weak
    direction (kappa ~ 60-100, as a chain's factors develop them) stay out
    of the queue: the probe estimate works on the Jacobi-scaled matrix, for
    which such rows read ~K (an estimate with max_i P_ii in place of the
    scaling reads 700-1100 here, above the threshold of 500).  Without a
    queued row the refined sampler's output equals the unrefined one bit for
    bit, and both are within 1e-4 of the oracle."""
    announce("spread component scales K=%d" % K)
    rng = np.random.default_rng(K + 7)
    ncol = 30000
    scale = np.exp(np.linspace(np.log(lo), np.log(hi), K))
    V = f32(rng.standard_normal((ncol, K)) * scale * 0.5)
    degs = [2000, 4000, 8000, 27000, 300, 150]
    r = np.concatenate([np.full(d, i) for i, d in enumerate(degs)])
    c = np.concatenate([rng.choice(ncol, d, replace=False) for d in degs])
    y = np.sum(rng.standard_normal((len(degs), K))[r] * V[c], 1) + 0.3 * rng.standard_normal(len(r))
    modes, mean, _ = O.matrix_rows(r, c, y, len(degs), ncol)
    Lam, mu = spd(rng, K, 1.0), np.zeros(K)
    ref = O.sample_latents(K, modes[0], mean, [V], 10.0, mu, Lam, SEED, 3, 0)
    a = gpu_sample_latents(K, modes[0], mean, [V], 10.0, mu, Lam, SEED, 3, 0, _lib.SAMPLER_BATCHED)
    b = gpu_sample_latents(K, modes[0], mean, [V], 10.0, mu, Lam, SEED, 3, 0, _lib.SAMPLER_BATCHED_NOREFINE)
    np.testing.assert_array_equal(a, b)
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("K=%d spread scales: row errors %s" % (K, " ".join("%.1e" % e for e in err)))
    assert err.max() <= 1e-4


def test_refinement_counts_and_switch():
    announce("refinement counters in a model")
    L = _lib.lib()
    K = 32
    rows, mean, V, Lam, mu, degs = _collinear_case(K, 5)
    r = np.repeat(np.arange(rows.nrows), np.diff(rows.row_ptr))
    c = rows.idx[:, 0]
    y = rows.vals
    stats = {}
    for on in (1, 0):
        m = GpuModel(2, K, 10.0)
        m.data(np.stack([r, c], 1), y, [rows.nrows, V.shape[0]])
        m.prior()
        m.prior()
        check(L.macau_set_refine(m.h, on, 1.0001))  # a threshold every heavy row exceeds
        m.setup(1, 2, SEED, _lib.SAMPLER_BATCHED)
        m.iterate(3)
        f, d = ctypes.c_longlong(), ctypes.c_longlong()
        check(L.macau_get_refine_stats(m.h, ctypes.byref(f), ctypes.byref(d)))
        stats[on] = (f.value, d.value)
        print("refinement %s: %d rows flagged, %d refined over 3 iterations" % ("on" if on else "off", f.value,
                                                                                 d.value))
    assert stats[0] == (0, 0)
    assert stats[1][0] > 0 and stats[1][1] == stats[1][0]


def test_refinement_of_many_rows_matches_oracle():
    """More queued rows than half the refinement grid (300 rows of 150-500
    nonzeros, all queued by a threshold every row exceeds): one work-group
    per row.  The model's iterations 1-2 stay within 1e-4 of the FP64 oracle
    and every queued row is counted as refined."""
    announce("refinement of 300 rows per half-sweep")
    L = _lib.lib()
    K = 32
    rng = np.random.default_rng(11)
    nrow, ncol = 300, 3000
    deg = rng.integers(150, 500, nrow)
    r = np.repeat(np.arange(nrow), deg)
    c = np.concatenate([rng.choice(ncol, d, replace=False) for d in deg])
    U, V = rng.standard_normal((nrow, 6)), rng.standard_normal((ncol, 6))
    y = np.sum(U[r] * V[c], 1) + 0.4 * rng.standard_normal(len(r))
    rt, ct = rng.integers(0, nrow, 500), rng.integers(0, ncol, 500)
    yt = np.sum(U[rt] * V[ct], 1)
    m = GpuModel(2, K, 4.0)
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([rt, ct], 1), yt)
    m.prior()
    m.prior()
    check(L.macau_set_refine(m.h, 1, 1.0001))  # every probed row is queued
    m.setup(1, 2, SEED, _lib.SAMPLER_BATCHED)
    modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
    ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
    for n_it in (1, 2):
        m.iterate(1)
        ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, precision=4.0, burnin=1, nsamples=2, seed=SEED,
                    max_iters=n_it, want_U=True)
        e0 = row_rel_err(m.latents(0, nrow), ref.U[0])
        e1 = row_rel_err(m.latents(1, ncol), ref.U[1])
        print("iteration %d: row errors %.2e (rows) %.2e (columns)" % (n_it, e0, e1))
        assert e0 < 1e-4 and e1 < 1e-4
    f, d = ctypes.c_longlong(), ctypes.c_longlong()
    check(L.macau_get_refine_stats(m.h, ctypes.byref(f), ctypes.byref(d)))
    assert f.value == d.value == 2 * nrow  # the rows mode, two iterations (columns have < 128 nonzeros)
