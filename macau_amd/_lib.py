"""ctypes binding of libmacau_amd.so (the C ABI declared in include/macau_amd.h).

The library is built in-tree by ``make -C macau_amd/csrc`` (or
``__graft_entry__.build()``).  There is no CPU fallback: importing a function
that needs the library raises if it is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmacau_amd.so")
_lib = None

SAMPLER_PLAIN_FP32 = 0
SAMPLER_BATCHED = 1
SAMPLER_BATCHED_NOREFINE = 2  # macau_sample_latents only

_p = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong
_ull = ctypes.c_ulonglong
_u = ctypes.c_uint

# name -> (restype, argtypes); every symbol declared in include/macau_amd.h
SIGNATURES = {
    "macau_last_error": (ctypes.c_char_p, []),
    "macau_version": (ctypes.c_char_p, []),
    "macau_make_fixed": (_p, [_i, _i, _d]),
    "macau_make_adaptive": (_p, [_i, _i, _d, _d]),
    "macau_make_probit": (_p, [_i, _i]),
    "macau_free": (None, [_p]),
    "macau_add_bpmf_prior": (_i, [_p]),
    "macau_add_macau_prior": (_i, [_p, _i, _i, _ll, _p, _p, _p, _d, _d]),
    "macau_add_macau_prior_ff": (_i, [_p, _i, _i, _ll, _p, _p, _p, _d, _d, _i]),
    "macau_add_macau_one_prior": (_i, [_p, _i, _i, _ll, _p, _p, _p, _d]),
    "macau_set_samples": (_i, [_p, _i, _i]),
    "macau_set_relation_data": (_i, [_p, _p, _i, _p, _i, _p]),
    "macau_set_relation_data_test": (_i, [_p, _p, _i, _p, _i, _p]),
    "macau_set_verbose": (_i, [_p, _i]),
    "macau_set_save_model": (_i, [_p, _i]),
    "macau_set_save_prefix": (_i, [_p, ctypes.c_char_p]),
    "macau_run": (_i, [_p]),
    "macau_get_rmse_test": (_d, [_p]),
    "macau_get_test_nonzeros": (_ll, [_p]),
    "macau_get_predictions": (_i, [_p, _p]),
    "macau_get_stds": (_i, [_p, _p]),
    "macau_get_test_data": (_i, [_p, _p]),
    "macau_set_seed": (_i, [_p, _ull]),
    "macau_set_sampler": (_i, [_p, _i]),
    "macau_sampler_supported": (_i, [_i, _i, _i]),
    "macau_set_graph": (_i, [_p, _i]),
    "macau_init": (_i, [_p]),
    "macau_iterate": (_i, [_p, _i]),
    "macau_synchronize": (_i, [_p]),
    "macau_get_iteration": (_i, [_p]),
    "macau_profile": (_i, [_p, _i, _p]),
    "macau_get_rmse_onesample": (_d, [_p]),
    "macau_get_rmse_avg": (_d, [_p]),
    "macau_get_auc": (_i, [_p, _p, _p]),
    "macau_get_noise_precision": (_d, [_p]),
    "macau_get_mean_value": (_d, [_p]),
    "macau_get_dims": (_i, [_p, _p]),
    "macau_get_latents": (_i, [_p, _i, _p]),
    "macau_get_hyper": (_i, [_p, _i, _p, _p]),
    "macau_get_link": (_i, [_p, _i, _p]),
    "macau_get_link_lambda": (_d, [_p, _i]),
    "macau_get_last_cg_iters": (_i, [_p, _i]),
    "macau_get_ftf_solver": (_i, [_p, _i]),
    "macau_set_refine": (_i, [_p, _i, _d]),
    "macau_get_refine_stats": (_i, [_p, _p, _p]),
    "macau_autotune": (_i, [_p, _i, ctypes.c_char_p, _p, _p]),
    "macau_get_tuning": (_i, [_p, _i, _p]),
    "macau_batch_create": (_p, [_p, _i]),
    "macau_batch_free": (None, [_p]),
    "macau_batch_set_graph": (_i, [_p, _i]),
    "macau_batch_iterate": (_i, [_p, _i]),
    "macau_batch_synchronize": (_i, [_p]),
    "macau_sample_latents": (_i, [_i, _i, _i, _p, _p, _p, _d, _p, _p, _d, _p, _p, _p, _ull, _i, _i, _i, _p]),
    "macau_cond_normal_wishart": (_i, [_i, _i, _p, _p, _d, _p, _d, _ull, _i, _i, _p, _p]),
    "macau_cond_normal_wishart_noalloc": (_i, [_i, _i, _p, _p, _d, _p, _d, _ull, _i, _i, _p, _p, _p, _p]),
    "macau_solve_blockcg": (_i, [_i, _i, _i, _ll, _p, _p, _p, _d, _p, _d, _p, _p]),
    "macau_rng_probe": (_i, [_ull, _u, _u, _u, _i, _p, _p, _p]),
    "macau_gamma_probe": (_i, [_ull, _p, _i, _u, _u, _p]),
    "macau_truncnorm_probe": (_i, [_ull, _u, _u, _u, _i, _p, _p, _p, _p]),
}


def lib():
    """Load the native library, failing loudly if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("macau_amd native library is missing (%s); build it with "
                               "`make -C macau_amd/csrc`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().macau_last_error().decode()


def check(rc):
    if rc != 0:
        raise RuntimeError(last_error())
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
