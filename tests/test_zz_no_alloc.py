"""The allocation-free Normal-Wishart draw against the allocating one.

macau_cond_normal_wishart_noalloc takes device pointers and keeps its
working memory in one static device array; two calls in a row (the second
with fewer columns and a smaller latent size, so it runs on what the first
left in that array) must give the draws of macau_cond_normal_wishart for the
same inputs and keys.  Device memory for the inputs and outputs comes from
the HIP runtime the library itself links (hipMalloc through ctypes).
"""
import ctypes

import numpy as np
import pytest

from _gpu_helpers import SEED, announce, gpu_cond_normal_wishart, spd
from macau_amd import _lib
from macau_amd._lib import ptr

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        _lib.lib()  # the library's own libamdhip64 is then the one loaded by soname
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def ok(self, rc, what):
        assert rc == 0, "%s failed (%d)" % (what, rc)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.bufs.append(p)
        self.ok(self.h.hipMemcpy(p, ptr(a), a.nbytes, 1), "hipMemcpy H2D")
        return p

    def empty(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(p), max(nbytes, 8)), "hipMalloc")
        self.bufs.append(p)
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        self.ok(self.h.hipMemcpy(ptr(out), p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc_draw(hip, U, mu0, T, kappa, nu, it, mode):
    N, K = U.shape  # U as (N, K) C-order = K x N column-major
    dU, dmu0, dT = hip.upload(U), hip.upload(mu0), hip.upload(np.asfortranarray(T).ravel(order="F"))
    dmu, dL, dst = hip.empty(8 * K), hip.empty(8 * K * K), hip.empty(4)
    rc = _lib.lib().macau_cond_normal_wishart_noalloc(K, N, dU, dmu0, kappa, dT, nu, SEED, it, mode, dmu, dL, dst,
                                                      None)
    assert rc == 0, _lib.last_error()
    hip.ok(hip.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
    status = int(hip.download(dst, 1, np.int32)[0])
    mu = hip.download(dmu, K, np.float64)
    Lam = hip.download(dL, K * K, np.float64).reshape(K, K, order="F")
    return mu, Lam, status


def test_noalloc_normal_wishart_matches_allocating_entry():
    hip = Hip()
    try:
        rng = np.random.default_rng(5)
        for K, N, it in ((64, 20000, 3), (10, 700, 4)):
            announce("allocation-free Normal-Wishart K=%d N=%d" % (K, N))
            U = rng.standard_normal((N, K)) * 0.5 + 0.1
            T = spd(rng, K, 0.7)
            mu0 = rng.standard_normal(K) * 0.1
            mu, Lam, status = _noalloc_draw(hip, U, mu0, T, 2.0, K + 3.0, it, 1)
            mur, Lamr = gpu_cond_normal_wishart(U, mu0, 2.0, T, K + 3.0, SEED, it, 1)
            e_mu = np.linalg.norm(mu - mur) / np.linalg.norm(mur)
            e_L = np.linalg.norm(Lam - Lamr) / np.linalg.norm(Lamr)
            print("allocation-free draw K=%d N=%d: mu rel err %.2e, Lambda rel err %.2e" % (K, N, e_mu, e_L))
            assert status == 0
            assert e_mu < 1e-10 and e_L < 1e-10
    finally:
        hip.close()


def test_noalloc_rejects_sizes_beyond_its_array():
    announce("allocation-free Normal-Wishart size limit")
    rc = _lib.lib().macau_cond_normal_wishart_noalloc(65, 10, None, None, 2.0, None, 68.0, SEED, 0, 0, None, None,
                                                      None, None)
    assert rc != 0
    assert "exceeds the allocation-free limit of 64" in _lib.last_error()
