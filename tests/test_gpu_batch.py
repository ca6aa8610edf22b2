"""Many independent BPMF instances as one launch set (macau_batch_*).

A scaled-down form of the 64-instance uneven workload (macau_amd.synthetic.c5
with every row and column count shrunk): instances of num_latent 8, 16, 32,
64 and 128, each checked against its own FP64 oracle run with the same data
and seed (every latent row within 1e-4 after iterations 1 and 2, the
posterior-mean test RMSE of a 30-iteration chain within 0.5%), and the batch
against the same models advanced one by one (identical chains: the models
keep their own state and draws).
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from _gpu_helpers import SEED, GpuModel, announce, row_rel_err
from macau_amd import _lib, synthetic
from macau_amd._lib import check

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _instances():
    inst = synthetic.c5(seed=5, n_instances=12, scale=0.08)
    Ks = [s["num_latent"] for _, _, _, s in inst]
    assert set(Ks) == {8, 16, 32, 64, 128}, Ks
    return inst


def _model(inst, sampler):
    (r, c, y), (rt, ct, yt), (nrow, ncol), st = inst
    m = GpuModel(2, st["num_latent"], st["precision"])
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([rt, ct], 1), yt)
    m.prior()
    m.prior()
    m.setup(3, 27, SEED, sampler)
    return m


def _sampler_for(K):
    L = _lib.lib()
    return _lib.SAMPLER_BATCHED if L.macau_sampler_supported(_lib.SAMPLER_BATCHED, K, 2) else \
        _lib.SAMPLER_PLAIN_FP32


def _batch(models):
    L = _lib.lib()
    arr = (ctypes.c_void_p * len(models))(*[m.h for m in models])
    b = L.macau_batch_create(arr, len(models))
    assert b, _lib.last_error()
    return b


def test_batch_instances_match_their_oracle_runs():
    L = _lib.lib()
    inst = _instances()
    models = [_model(x, _sampler_for(x[3]["num_latent"])) for x in inst]
    b = _batch(models)
    try:
        refs = []
        for q, ((r, c, y), (rt, ct, yt), (nrow, ncol), st) in enumerate(inst):
            modes, mean, nnz = O.matrix_rows(r, c, y, nrow, ncol)
            ti, tv = O.matrix_test(rt, ct, yt, nrow, ncol)
            refs.append((modes, mean, nnz, ti, tv, nrow, ncol, st))
        for it in (1, 2):
            announce("batch of %d instances, iteration %d" % (len(models), it))
            check(L.macau_batch_iterate(b, 1))
            check(L.macau_batch_synchronize(b))
            for q, (modes, mean, nnz, ti, tv, nrow, ncol, st) in enumerate(refs):
                K = st["num_latent"]
                ref = O.run(modes, mean, nnz, [nrow, ncol], K, ti, tv, precision=st["precision"], burnin=3,
                            nsamples=27, seed=SEED, max_iters=it, want_U=True)
                e0 = row_rel_err(models[q].latents(0, nrow), ref.U[0])
                e1 = row_rel_err(models[q].latents(1, ncol), ref.U[1])
                print("instance %d (%dx%d, K=%d, %d train): iteration %d row err %.2e / %.2e"
                      % (q, nrow, ncol, K, nnz, it, e0, e1))
                assert e0 < TOL and e1 < TOL
        announce("batch chains to 30 iterations")
        check(L.macau_batch_iterate(b, 28))
        check(L.macau_batch_synchronize(b))
        for q, (modes, mean, nnz, ti, tv, nrow, ncol, st) in enumerate(refs):
            ref = O.run(modes, mean, nnz, [nrow, ncol], st["num_latent"], ti, tv, precision=st["precision"],
                        burnin=3, nsamples=27, seed=SEED)
            d = abs(models[q].rmse() - ref.rmse) / ref.rmse
            print("instance %d: posterior-mean RMSE %.6f vs oracle %.6f (%.1e)" % (q, models[q].rmse(), ref.rmse, d))
            assert d < 0.005
    finally:
        L.macau_batch_free(b)


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_equals_models_advanced_alone(graph):
    L = _lib.lib()
    inst = _instances()[:6]
    announce("batch vs alone graph=%d" % graph)
    together = [_model(x, _sampler_for(x[3]["num_latent"])) for x in inst]
    alone = [_model(x, _sampler_for(x[3]["num_latent"])) for x in inst]
    b = _batch(together)
    try:
        check(L.macau_batch_set_graph(b, graph))
        check(L.macau_batch_iterate(b, 4))
        check(L.macau_batch_synchronize(b))
        for m in alone:
            m.iterate(4)
        for x, mt, ma in zip(inst, together, alone):
            nrow, ncol = x[2]
            assert L.macau_get_iteration(mt.h) == L.macau_get_iteration(ma.h) == 4
            np.testing.assert_array_equal(mt.latents(0, nrow), ma.latents(0, nrow))
            np.testing.assert_array_equal(mt.latents(1, ncol), ma.latents(1, ncol))
            assert mt.rmse() == ma.rmse()
    finally:
        L.macau_batch_free(b)


def test_batch_rejects_uninitialised_model():
    announce("batch of an uninitialised model")
    L = _lib.lib()
    h = L.macau_make_fixed(2, 4, 1.0)
    try:
        arr = (ctypes.c_void_p * 1)(h)
        assert not L.macau_batch_create(arr, 1)
        assert "initialised" in _lib.last_error()
    finally:
        L.macau_free(h)
