"""Autotuner of the batched sampler's launch shape (macau_autotune).

Tuning times candidate shapes on the device and must leave the chain where it
was; the chain continued with the tuned shape equals the untuned chain (the
shape only reorders FP32 sums of heavy-row segments and FP64 sums of SSE
partials); the recorded choice is found again at macau_init through
MACAU_AMD_TUNE_CACHE.
"""
import numpy as np
import pytest

from _gpu_helpers import SEED, GpuModel, announce, power_law_rows, row_rel_err
from macau_amd import _lib
from macau_amd._lib import check, ptr

pytestmark = pytest.mark.gpu


def _data():
    rng = np.random.default_rng(77)
    nrow, ncol = 2500, 2200
    r, c, y = power_law_rows(rng, nrow, ncol, 60000, heavy=(2100, 1900, 3000))
    y = np.round(y * 2) / 2
    return r, c, y, nrow, ncol


def _model(K, r, c, y, nrow, ncol):
    m = GpuModel(2, K, "adaptive")
    m.data(np.stack([r, c], 1), y, [nrow, ncol], np.stack([r[:500], c[:500]], 1), y[:500])
    m.prior()
    m.prior()
    m.setup(2, 6, SEED, _lib.SAMPLER_BATCHED)
    return m


def _knobs(m, mode):
    k = np.zeros(4, dtype=np.int32)
    tuned = _lib.lib().macau_get_tuning(m.h, mode, ptr(k))
    assert tuned in (0, 1), _lib.last_error()
    return tuned, k.tolist()


@pytest.mark.parametrize("K", [32, 64])
def test_autotune_keeps_the_chain_and_records_the_choice(K, tmp_path, monkeypatch):
    announce("autotune K=%d" % K)
    L = _lib.lib()
    r, c, y, nrow, ncol = _data()
    monkeypatch.setenv("MACAU_AMD_TUNE_DISABLE", "1")
    a = _model(K, r, c, y, nrow, ncol)
    b = _model(K, r, c, y, nrow, ncol)
    assert _knobs(a, 0) == (0, [2048, 8192, 2048, 16384])
    a.iterate(2)
    b.iterate(2)
    before = [a.latents(0, nrow), a.latents(1, ncol)]
    rec = str(tmp_path / "tune.txt")
    dms, tms = np.zeros(2), np.zeros(2)
    check(L.macau_autotune(a.h, 2, rec.encode(), ptr(dms), ptr(tms)))
    print("autotune K=%d: default ms %s, tuned ms %s, shapes %s %s" % (K, dms, tms, _knobs(a, 0), _knobs(a, 1)))
    assert np.all(dms > 0) and np.all(tms > 0) and np.all(tms <= dms)
    np.testing.assert_array_equal(a.latents(0, nrow), before[0])
    np.testing.assert_array_equal(a.latents(1, ncol), before[1])
    assert L.macau_get_iteration(a.h) == 2
    a.iterate(3)
    b.iterate(3)
    for mode, n in ((0, nrow), (1, ncol)):
        assert row_rel_err(a.latents(mode, n), b.latents(mode, n)) < 1e-5
    assert abs(a.alpha() - b.alpha()) <= 1e-6 * b.alpha()
    lines = [ln for ln in open(rec).read().splitlines() if ln and not ln.startswith("#")]
    assert len(lines) == 2
    # a new model of the same data finds the recorded shapes at init
    monkeypatch.delenv("MACAU_AMD_TUNE_DISABLE")
    monkeypatch.setenv("MACAU_AMD_TUNE_CACHE", rec)
    d = _model(K, r, c, y, nrow, ncol)
    for mode in (0, 1):
        assert _knobs(d, mode) == (1, _knobs(a, mode)[1])
    d.iterate(5)
    assert row_rel_err(d.latents(0, nrow), b.latents(0, nrow)) < 1e-5
