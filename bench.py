#!/usr/bin/env python3
"""Benchmark: Gibbs iterations/s of the BPMF/Macau sampler on one MI355X.

Default workload (the headline line): the 138,000 x 27,000 recommender-scale
BPMF with 20M observed cells (18M train, 2M test), K = 32, adaptive noise
precision (macau_amd.synthetic.c2).  A step is one full Gibbs iteration (both
half-sweeps, both Normal-Wishart draws, the noise update and the test
evaluation) with all inputs resident in HBM.  The headline value is the
batched mixed-precision sampler; the plain FP32 sampler is timed beside it on
the same data, and the posterior-mean test RMSE of both full chains
(burnin 50 + 100 samples) is reported.

The dominant kernel group is the per-row latent sampler (both half-sweeps);
its duration is measured with HIP events on the model's stream
(macau_profile) and priced against the MFMA/vector peak of the arithmetic it
uses: algorithmic FLOPs per half-sweep = nnz (K^2 + 3K) + rows (K^3/3 + 4K^2).

cpu_baseline: the FP64 C/OpenMP oracle timed on this host for a bounded
sample of the same workload (25 full Gibbs iterations, about 10 s on 16
cores), with the host CPU model.

RMSE gate (headline workload): the oracle's full FP64 chain on the same data
and seed runs on the host cores while the GPU chains run (after the timed
region); the line reports each GPU chain's posterior-mean test RMSE relative
to it, and every sampled quantity's error against the oracle after two
iterations from the same initial state (rows and columns latents, mu and
Lambda of both modes, alpha).

Other workloads (`--workload`, one JSON line each, not the headline):
  c3  Macau with 100,000-dim sparse binary row features (block CG for the
      link matrix every iteration); reports the CG iterations per solve and
      the achieved HBM bandwidth of the row prior update;
  c3f the same data with 4,000-dim row features, for which the link matrix
      takes the F^T F branch (side information with at most 10,000 features,
      as the Python API chooses): one eigendecomposition of F^T F at setup,
      then two FP64 products per iteration; reports their HBM rate;
  c4  3-mode tensor 50,000 x 2,000 x 100, 10M cells, K = 64;
  c3u the c3 data with the univariate prior (MacauOnePrior) on the rows:
      coordinate-wise row updates and, per latent, a sequential sweep over
      the 100,000 features for the link matrix;
  c2p the c2 cells binarised at the train mean, probit noise (the
      augmentation kernel draws one truncated normal per observation before
      each half-sweep); the full-chain line reports test AUC instead of RMSE;
  c5  64 independent BPMF instances of uneven size and num_latent in
      {8, 16, 32, 64, 128} advanced as one launch set (macau_batch_*: one HIP
      graph per Gibbs iteration of all instances); a step is one iteration of
      every instance; also reports the same models advanced one after another.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from macau_amd import _lib, synthetic  # noqa: E402
from macau_amd._lib import check, ptr  # noqa: E402

METRIC = "Gibbs iters/sec, batched-mixed vs plain FP32, 20M nnz K=32, RMSE≤+0.5%"
PEAK_TFLOPS = {"plain": 157.3, "batched": 2500.0, "batched-norefine": 2500.0}  # FP32 vector/MFMA; BF16 dense MFMA
UNIT_OF = {"plain": "f32", "batched": "bf16x4-f32acc", "batched-norefine": "bf16x4-f32acc"}
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
MAX_FF_SIZE = 10000  # side information with at most this many features uses F^T F (macau.pyx:334)


def load_workload(name, seed):
    """Generic N-mode form: idx [nnz x nmodes], vals, test idx/vals, dims,
    settings, side (per mode: scipy COO or None), label."""
    if name == "c1":
        train, test, dims, settings = synthetic.c1(seed)
        return dict(idx=np.stack([train[0], train[1]], 1), vals=train[2], tidx=np.stack([test[0], test[1]], 1),
                    tvals=test[2], dims=dims, settings=settings, side=[None, None],
                    label="BPMF 5000x2000, 500k observed (400k train / 100k test), uniform cells, K=10, "
                          "fixed precision 5, 20 burn-in + 50 samples")
    if name in ("c2", "c2p"):
        train, test, dims, settings = synthetic.c2(seed)
        idx = np.stack([train[0], train[1]], 1)
        tidx = np.stack([test[0], test[1]], 1)
        vals, tvals = train[2], test[2]
        if name == "c2":
            return dict(idx=idx, vals=vals, tidx=tidx, tvals=tvals, dims=dims, settings=settings,
                        side=[None, None],
                        label="BPMF 138000x27000, 20M observed (18M train / 2M test), Zipf-1.1 degrees, K=32, "
                              "adaptive noise")
        cut = float(np.mean(vals))
        settings = dict(settings, precision="probit")
        return dict(idx=idx, vals=(vals > cut).astype(np.float64), tidx=tidx,
                    tvals=(tvals > cut).astype(np.float64), dims=dims, settings=settings, side=[None, None],
                    label="BPMF probit 138000x27000, 20M observed (18M train / 2M test) binarised at the train "
                          "mean, Zipf-1.1 degrees, K=32")
    if name in ("c3", "c3f", "c3u"):
        nfeat = 4000 if name == "c3f" else 100000
        (idx, vals), (tidx, tvals), dims, settings, side = synthetic.c3(seed, nfeat=nfeat)
        how = "CG tol 1e-6" if nfeat > MAX_FF_SIZE else "F^T F branch"
        if name == "c3u":
            settings = dict(settings, univariate=True)
            how = "univariate prior (coordinate-wise rows, feature-sequential link matrix)"
        return dict(idx=idx, vals=vals, tidx=tidx, tvals=tvals, dims=dims, settings=settings, side=side,
                    label="Macau 15000x350, 60k observed, %d-dim binary row features (~%d nnz), K=32, "
                          "lambda_beta 5, %s" % (nfeat, side[0].nnz, how))
    if name == "c4":
        (idx, vals), (tidx, tvals), dims, settings, side = synthetic.c4(seed)
        return dict(idx=idx, vals=vals, tidx=tidx, tvals=tvals, dims=dims, settings=settings, side=side,
                    label="BPMF tensor 50000x2000x100, 10M observed (9M train / 1M test), Zipf-1.2 degrees, "
                          "K=64, fixed precision 11.1")
    raise ValueError(name)


def build_model(w, sampler_kind, seed, graph=True, refine=True):
    L = _lib.lib()
    st = w["settings"]
    K = st["num_latent"]
    nm = len(w["dims"])
    if st["precision"] == "adaptive":
        h = L.macau_make_adaptive(nm, K, 1.0, 10.0)
    elif st["precision"] == "probit":
        h = L.macau_make_probit(nm, K)
    else:
        h = L.macau_make_fixed(nm, K, float(st["precision"]))
    if not h:
        raise RuntimeError(_lib.last_error())
    dims = np.array(w["dims"], dtype=np.int32)
    idx = np.asfortranarray(w["idx"].astype(np.int32))
    check(L.macau_set_relation_data(h, ptr(idx), nm, ptr(np.ascontiguousarray(w["vals"], dtype=np.float64)),
                                    idx.shape[0], ptr(dims)))
    tidx = np.asfortranarray(w["tidx"].astype(np.int32))
    check(L.macau_set_relation_data_test(h, ptr(tidx), nm, ptr(np.ascontiguousarray(w["tvals"], dtype=np.float64)),
                                         tidx.shape[0], ptr(dims)))
    for s in w["side"]:
        if s is None:
            check(L.macau_add_bpmf_prior(h))
        else:
            r = np.ascontiguousarray(s.row, dtype=np.int32)
            c = np.ascontiguousarray(s.col, dtype=np.int32)
            v = np.ascontiguousarray(s.data, dtype=np.float64)
            if st.get("univariate"):
                check(L.macau_add_macau_one_prior(h, s.shape[0], s.shape[1], len(v), ptr(r), ptr(c), ptr(v),
                                                  float(st["lambda_beta"])))
            else:
                check(L.macau_add_macau_prior_ff(h, s.shape[0], s.shape[1], len(v), ptr(r), ptr(c), ptr(v),
                                                 float(st["lambda_beta"]), float(st["tol"]), MAX_FF_SIZE))
    check(L.macau_set_samples(h, st["burnin"], st["nsamples"]))
    check(L.macau_set_seed(h, seed))
    check(L.macau_set_sampler(h, sampler_kind))
    check(L.macau_set_refine(h, int(refine), 0.0))
    check(L.macau_set_graph(h, int(graph)))
    check(L.macau_set_verbose(h, 0))
    check(L.macau_init(h))
    return h


def sampler_flops(w):
    """Algorithmic FLOPs of the per-row samplers in one Gibbs iteration."""
    K = w["settings"]["num_latent"]
    nnz, nm = w["idx"].shape
    per_nnz = K * K + 3 * K + (nm - 2) * K  # Gram + RHS (+ the Hadamard gather for tensors)
    per_row = K ** 3 / 3.0 + 4 * K * K
    return nm * nnz * per_nnz + sum(w["dims"]) * per_row


def kind_of(name):
    return _lib.SAMPLER_PLAIN_FP32 if name == "plain" else _lib.SAMPLER_BATCHED


def time_sampler(name, w, steps, warmup, seed, log):
    L = _lib.lib()
    ts = time.perf_counter()
    h = build_model(w, kind_of(name), seed, refine=name != "batched-norefine")
    setup_s = time.perf_counter() - ts
    nm = len(w["dims"])
    try:
        check(L.macau_iterate(h, warmup))
        check(L.macau_synchronize(h))
        t0 = time.perf_counter()
        check(L.macau_iterate(h, steps))
        check(L.macau_synchronize(h))
        t1 = time.perf_counter()
        nprof = max(3, min(steps, 10))
        phase = np.zeros(2 * nm + 2)
        check(L.macau_profile(h, nprof, ptr(phase)))
        cg = [int(L.macau_get_last_cg_iters(h, m)) for m in range(nm)]
        cg.append(int(L.macau_get_ftf_solver(h, 0)))  # F^T F solver of the rows
        flagged, refined = ctypes.c_longlong(), ctypes.c_longlong()
        check(L.macau_get_refine_stats(h, ctypes.byref(flagged), ctypes.byref(refined)))
        its = warmup + steps + nprof
        shapes = []
        if name != "plain":
            for m in range(nm):
                k = np.zeros(4, dtype=np.int32)
                tuned = int(L.macau_get_tuning(h, m, ptr(k)))
                shapes.append({"seg_len": int(k[0]), "light_cap": int(k[1]), "heavy_cap": int(k[2]),
                               "seg_cap": int(k[3]), "from_tuning_table": bool(tuned == 1)})
        cg.append({"rows_flagged_per_iteration": flagged.value / its, "rows_refined_per_iteration": refined.value / its,
                   "launch_shapes": shapes})
        log("%s: %d steps in %.3f s -> %.1f it/s; phases(ms) %s" %
            (name, steps, t1 - t0, steps / (t1 - t0), np.round(phase, 3).tolist()))
    finally:
        L.macau_free(h)
    return (t1 - t0) / steps, phase, cg, setup_s


def full_chain_rmse(name, w, seed):
    L = _lib.lib()
    h = build_model(w, kind_of(name), seed, refine=name != "batched-norefine")
    try:
        n = w["settings"]["burnin"] + w["settings"]["nsamples"]
        t0 = time.perf_counter()
        check(L.macau_iterate(h, n))
        check(L.macau_synchronize(h))
        t1 = time.perf_counter()
        if w["settings"]["precision"] == "probit":
            a1, aa = ctypes.c_double(), ctypes.c_double()
            check(L.macau_get_auc(h, ctypes.byref(a1), ctypes.byref(aa)))
            return aa.value, n / (t1 - t0)
        return L.macau_get_rmse_avg(h), n / (t1 - t0)
    finally:
        L.macau_free(h)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Host cores this process may use (OMP_NUM_THREADS when set, else the
    affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


_ORACLE_INPUTS = {}


def oracle_inputs(w):
    """The FP64 oracle's view of a workload (train rows of every mode, test
    cells in the reference's order, side information as CSR pairs), built
    once."""
    import oracle as O  # the FP64 oracle: the CPU baseline and the parity checker
    key = id(w)
    if key not in _ORACLE_INPUTS:
        dims = w["dims"]
        if len(dims) == 2:
            r, c = w["idx"][:, 0], w["idx"][:, 1]
            modes, mean, nnz = O.matrix_rows(r, c, w["vals"], dims[0], dims[1])
            ti, tv = O.matrix_test(w["tidx"][:, 0], w["tidx"][:, 1], w["tvals"], dims[0], dims[1])
        else:
            modes, mean, nnz = O.tensor_rows(w["idx"], w["vals"], dims)
            ti, tv = O.tensor_test(w["tidx"], w["tvals"], dims)
        side = [None if s is None else O.csr_pair(s.row, s.col, s.data, s.shape[0], s.shape[1])
                for s in w["side"]]
        _ORACLE_INPUTS[key] = (modes, mean, nnz, ti, tv, side)
    return _ORACLE_INPUTS[key]


def oracle_run(w, seed, **kw):
    import oracle as O
    st = w["settings"]
    modes, mean, nnz, ti, tv, side = oracle_inputs(w)
    O.set_threads(host_threads())
    adaptive = st["precision"] == "adaptive"
    extra = {}
    if any(s is not None for s in side):
        extra = dict(side=side, lambda_beta=float(st["lambda_beta"]), tol=float(st["tol"]),
                     use_ftf=[s is not None and w["side"][m].shape[1] <= MAX_FF_SIZE for m, s in enumerate(side)],
                     uni=[s is not None and bool(st.get("univariate")) for s in side])
    return O.run(modes, mean, nnz, list(w["dims"]), st["num_latent"], ti, tv, is_tensor=len(w["dims"]) > 2,
                 adaptive=adaptive, precision=1.0 if adaptive else float(st["precision"]), burnin=st["burnin"],
                 nsamples=st["nsamples"], seed=seed, **extra, **kw)


def cpu_baseline(w, iters, seed, log):
    import oracle as O
    oracle_inputs(w)
    t0 = time.perf_counter()
    oracle_run(w, seed, max_iters=iters)
    dt = time.perf_counter() - t0
    threads = O.get_threads()
    log("cpu oracle: %d iterations in %.2f s on %d threads" % (iters, dt, threads))
    return {"value": iters / dt, "unit": "iters/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": "%d full Gibbs iterations of the same workload (%s), FP64, OpenMP dynamic,2 rows (C "
                      "restatement of the reference)" % (iters, w["label"])}


def fp64_errors(w, name, seed, iters, log):
    """Every sampled quantity of the GPU chain against the FP64 oracle after
    `iters` iterations from the same initial state and draws: the largest
    row-relative error of each mode's latents, the relative errors of each
    mode's mu and Lambda, and of the noise precision alpha."""
    L = _lib.lib()
    K = w["settings"]["num_latent"]
    dims = w["dims"]
    h = build_model(w, kind_of(name), seed, refine=name != "batched-norefine")
    try:
        check(L.macau_iterate(h, iters))
        check(L.macau_synchronize(h))
        out = {}
        ref = oracle_run(w, seed, max_iters=iters, want_U=True)
        for m, tag in enumerate(("rows", "cols")):
            U = np.zeros((dims[m], K))
            check(L.macau_get_latents(h, m, ptr(U)))
            Ur = ref.U[m]
            out["U_%s_max_row_rel" % tag] = float(np.max(np.linalg.norm(U - Ur, axis=1) /
                                                         np.maximum(np.linalg.norm(Ur, axis=1), 1e-300)))
            mu = np.zeros(K)
            Lam = np.zeros((K, K), order="F")
            check(L.macau_get_hyper(h, m, ptr(mu), ptr(Lam)))
            out["mu_%s_rel" % tag] = float(np.linalg.norm(mu - ref.mu[m]) / np.linalg.norm(ref.mu[m]))
            out["Lambda_%s_rel" % tag] = float(np.linalg.norm(Lam - ref.Lambda[m]) / np.linalg.norm(ref.Lambda[m]))
        out["alpha_rel"] = float(abs(L.macau_get_noise_precision(h) - ref.alpha) / ref.alpha)
        log("%s vs FP64 after %d iterations: %s" % (name, iters, json.dumps({k: "%.2e" % v for k, v in out.items()})))
        return out
    finally:
        L.macau_free(h)


def load_traffic(name):
    """HBM bytes of one iteration's sampler launches, from the FETCH_SIZE /
    WRITE_SIZE counter runs recorded in profiles/traffic.json (None if absent)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    e = d.get(name)
    return None if e is None else e.get("bytes_per_iteration")


def cg_bytes_per_iteration(w):
    """Algorithmic HBM bytes of one block-CG iteration (linop.h:279-332) for
    the row side information: F and F^T read once each by the two SpMMs
    (int32 index + FP64 value per nonzero), and 13 streaming passes over an
    nrhs x nfeat FP64 block (KP written, P read by the SpMM gather and the
    P^T KP product, KP read twice, X and R read and written, R read for
    R R^T, P read and written by the direction update, reg P read)."""
    F = w["side"][0]
    K = w["settings"]["num_latent"]
    return 2 * F.nnz * 12 + 13 * K * F.shape[1] * 8


def autotune(w, args, log):
    """Default launch shapes timed, then the autotuner (macau_autotune), whose
    choices go to a table that the models built afterwards read at init
    (MACAU_AMD_TUNE_CACHE)."""
    L = _lib.lib()
    os.environ["MACAU_AMD_TUNE_DISABLE"] = "1"
    sec_def, _, _, _ = time_sampler("batched", w, args.steps, args.warmup, args.seed, log)
    del os.environ["MACAU_AMD_TUNE_DISABLE"]
    rec = os.path.join(ROOT, "gpurun_out", "tuning_%s.txt" % args.workload)
    os.makedirs(os.path.dirname(rec), exist_ok=True)
    if os.path.exists(rec):
        os.remove(rec)
    h = build_model(w, kind_of("batched"), args.seed)
    nm = len(w["dims"])
    dms, tms = np.zeros(nm), np.zeros(nm)
    try:
        t0 = time.perf_counter()
        check(L.macau_autotune(h, 5, rec.encode(), ptr(dms), ptr(tms)))
        tune_s = time.perf_counter() - t0
        knobs = []
        for m in range(nm):
            k = np.zeros(4, dtype=np.int32)
            L.macau_get_tuning(h, m, ptr(k))
            knobs.append({"seg_len": int(k[0]), "light_cap": int(k[1]), "heavy_cap": int(k[2]), "seg_cap": int(k[3])})
    finally:
        L.macau_free(h)
    os.environ["MACAU_AMD_TUNE_CACHE"] = rec
    log("autotune (%.1f s): half-sweep ms default %s -> tuned %s; %s" %
        (tune_s, np.round(dms, 4).tolist(), np.round(tms, 4).tolist(), knobs))
    return {"default_iters_per_sec": 1.0 / sec_def, "tune_seconds": tune_s, "record": os.path.relpath(rec, ROOT),
            "modes": [{"default_ms": float(dms[m]), "tuned_ms": float(tms[m]), "shape": knobs[m]} for m in range(nm)]}


def c5_workloads(seed):
    out = []
    for q, (train, test, dims, st) in enumerate(synthetic.c5(seed)):
        out.append(dict(idx=np.stack([train[0], train[1]], 1), vals=train[2], tidx=np.stack([test[0], test[1]], 1),
                         tvals=test[2], dims=dims, settings=st, side=[None, None], label="instance %d" % q))
    return out


def bench_c5(args, log):
    """64 uneven instances as one launch set against the same models advanced
    one by one."""
    L = _lib.lib()
    t0 = time.perf_counter()
    ws = c5_workloads(args.seed)
    log("generated c5 in %.1f s: %d instances, %d train cells" % (time.perf_counter() - t0, len(ws),
                                                                   sum(len(w["vals"]) for w in ws)))
    kinds = []
    for w in ws:
        K = w["settings"]["num_latent"]
        kinds.append("batched" if L.macau_sampler_supported(_lib.SAMPLER_BATCHED, K, 2) else "plain")
    t0 = time.perf_counter()
    hs = [build_model(w, kind_of(k), args.seed + q) for q, (w, k) in enumerate(zip(ws, kinds))]
    log("built %d models in %.1f s" % (len(hs), time.perf_counter() - t0))
    arr = (ctypes.c_void_p * len(hs))(*hs)
    b = L.macau_batch_create(arr, len(hs))
    if not b:
        raise RuntimeError(_lib.last_error())
    try:
        check(L.macau_batch_iterate(b, args.warmup))
        check(L.macau_batch_synchronize(b))
        t0 = time.perf_counter()
        check(L.macau_batch_iterate(b, args.steps))
        check(L.macau_batch_synchronize(b))
        sec_graph = (time.perf_counter() - t0) / args.steps
        log("batch graph: %d steps in %.3f s -> %.1f batch iterations/s" %
            (args.steps, sec_graph * args.steps, 1.0 / sec_graph))
        # each model's own graph on its own streams, all queued before the wait
        check(L.macau_batch_set_graph(b, 0))
        check(L.macau_batch_iterate(b, 1))
        check(L.macau_batch_synchronize(b))
        t0 = time.perf_counter()
        check(L.macau_batch_iterate(b, args.steps))
        check(L.macau_batch_synchronize(b))
        sec_streams = (time.perf_counter() - t0) / args.steps
        log("batch streams: %d steps in %.3f s -> %.1f batch iterations/s" %
            (args.steps, sec_streams * args.steps, 1.0 / sec_streams))
        mode = "graph" if sec_graph <= sec_streams else "streams"
        sec = min(sec_graph, sec_streams)
        # the same models one after another (each its own graph), same number of iterations
        for h in hs:
            check(L.macau_iterate(h, 1))
            check(L.macau_synchronize(h))
        t0 = time.perf_counter()
        for h in hs:
            check(L.macau_iterate(h, args.steps))
            check(L.macau_synchronize(h))
        seq = (time.perf_counter() - t0) / args.steps
        log("one by one: %.3f ms per iteration of all instances" % (seq * 1e3))
        rmse = []
        if not args.no_chain:
            done = args.warmup + 3 * args.steps + 2
            total = max(w["settings"]["burnin"] + w["settings"]["nsamples"] for w in ws)
            check(L.macau_batch_iterate(b, max(total - done, 0)))
            check(L.macau_batch_synchronize(b))
            rmse = [float(L.macau_get_rmse_avg(h)) for h in hs]
            log("posterior-mean test RMSE after %d iterations: min %.4f median %.4f max %.4f (noise sd 0.5)" %
                (max(total, done), min(rmse), float(np.median(rmse)), max(rmse)))
    finally:
        L.macau_batch_free(b)
        for h in hs:
            L.macau_free(h)
    flops = {"plain": 0.0, "batched": 0.0}
    by_k = {}
    for w, k in zip(ws, kinds):
        flops[k] += sampler_flops(w)
        K = w["settings"]["num_latent"]
        e = by_k.setdefault(str(K), {"instances": 0, "train_cells": 0, "sampler": k})
        e["instances"] += 1
        e["train_cells"] += len(w["vals"])
    rows = sum(sum(w["dims"]) for w in ws)
    out = {
        "metric": "Gibbs iterations/sec of the 64-instance uneven batch (c5), every instance one iteration per step",
        "value": 1.0 / sec,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x4-f32acc" if all(k == "batched" for k in kinds) else "bf16x4-f32acc (batched) / f32 (plain)",
        "data": "synthetic",
        "config": {"workload": "64 independent BPMF instances, rows x cols log-uniform in [1000,50000] x "
                               "[500,10000], density log-uniform in [0.1%,5%], K in {8,16,32,64,128}, fixed "
                               "precision 4", "parallelism": "single-gpu", "global_batch": sum(len(w["vals"]) for w in ws)},
        "instance_iterations_per_sec": len(ws) / sec,
        "row_updates_per_sec": rows / sec,
        "launch": mode,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "graph_ms_per_step": sec_graph * 1e3,
        "streams_ms_per_step": sec_streams * 1e3,
        "one_by_one_ms_per_step": seq * 1e3,
        "speedup_vs_one_by_one": seq / sec,
        "instances_by_latent": by_k,
        "sampler_tflops": {k: v / sec / 1e12 for k, v in flops.items()},
        "note": "sampler_tflops = algorithmic sampler FLOPs of the instances on each sampler / whole batch step time",
    }
    if rmse:
        out["rmse"] = {"min": min(rmse), "median": float(np.median(rmse)), "max": max(rmse), "noise_sd": 0.5}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c3f", "c3u", "c4", "c2p", "c5"], default="c2")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 oracle chain and error table")
    ap.add_argument("--cpu-iters", type=int, default=25)
    ap.add_argument("--sampler", choices=["batched", "batched-norefine", "plain"], default=None,
                    help="batched-norefine: the batched sampler without the condition probes and refinement")
    ap.add_argument("--hw-queues", type=int, default=16, choices=range(1, 33), metavar="1..32",
                    help="hardware queues per process for --workload c5 (GPU_MAX_HW_QUEUES)")
    ap.add_argument("--tune", action="store_true",
                    help="time the default launch shapes, run the autotuner, time the tuned shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (one GPU per box)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    if args.workload == "c5":
        # 64 models with a graph and streams each: with 16 hardware queues per
        # process (HIP's default is 4) they serialise least (batch it/s with
        # 4, 8, 12, 16, 32 queues: 24.7, 23.8, 24.8, 27.2, 12.8); read when the
        # library first touches the GPU, so set before that
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
        return bench_c5(args, log)
    t0 = time.perf_counter()
    w = load_workload(args.workload, args.seed)
    st = w["settings"]
    K, nm = st["num_latent"], len(w["dims"])
    log("generated %s in %.1f s: %d train, %d test" % (args.workload, time.perf_counter() - t0, len(w["vals"]),
                                                        len(w["tvals"])))
    batched_ok = bool(_lib.lib().macau_sampler_supported(_lib.SAMPLER_BATCHED, K, nm))
    head = args.sampler or ("batched" if batched_ok else "plain")
    res = {}
    tuning = None
    if args.tune and head == "batched":
        tuning = autotune(w, args, log)
    names = [head] + ([] if (args.no_plain or head == "plain") else ["plain"])
    for name in names:
        sec, phase, cg, setup_s = time_sampler(name, w, args.steps, args.warmup, args.seed, log)
        res[name] = {"sec": sec, "phase": phase, "cg": cg, "setup_s": setup_s}
    # the headline workload's RMSE gate is against the FP64 oracle's chain on
    # the same data and seed; it runs on the host cores (a thread: the
    # oracle's C code runs without the GIL) while the GPU chains run, after
    # the timed region
    fp64 = {}
    gate = args.workload in ("c1", "c2") and not args.no_chain and not args.no_fp64
    if gate:
        import threading

        def fp64_chain():
            t0 = time.perf_counter()
            r = oracle_run(w, args.seed)
            fp64["rmse"], fp64["sec"] = r.rmse, time.perf_counter() - t0

        oracle_inputs(w)
        th = threading.Thread(target=fp64_chain)
        th.start()
    if not args.no_chain:
        for name in names:
            rmse, ips = full_chain_rmse(name, w, args.seed)
            res[name]["rmse"] = rmse
            log("%s full chain (%d iterations): posterior-mean %s %.6f" %
                (name, st["burnin"] + st["nsamples"], "AUC" if st["precision"] == "probit" else "RMSE", rmse))
    if gate:
        th.join()
        log("FP64 oracle full chain (%d iterations): posterior-mean RMSE %.6f in %.1f s" %
            (st["burnin"] + st["nsamples"], fp64["rmse"], fp64["sec"]))
        for name in names:
            res[name]["fp64_err"] = fp64_errors(w, name, args.seed, 2, log)

    flops = sampler_flops(w)
    ph = res[head]["phase"]
    sampler_ms = float(sum(ph[:nm]))
    achieved = flops / (sampler_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS[head],
            "traffic": load_traffic(head) if args.workload == "c2" else None,
            "kernel": "per-row latent sampler, all half-sweeps (%s)" % head,
            "flops_per_iteration": flops, "ms_per_iteration": sampler_ms}
    sec = res[head]["sec"]
    mode_names = ["rows", "cols"] if nm == 2 else ["mode%d" % m for m in range(nm)]
    phase_ms = {}
    for m in range(nm):
        phase_ms["sample_" + mode_names[m]] = float(ph[m])
    for m in range(nm):
        phase_ms["prior_" + mode_names[m]] = float(ph[nm + m])
    phase_ms["noise"] = float(ph[2 * nm])
    phase_ms["eval"] = float(ph[2 * nm + 1])
    out = {
        "metric": METRIC if args.workload == "c2" else "Gibbs iters/sec (%s)" % args.workload,
        "value": 1.0 / sec,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": UNIT_OF[head],
        "data": "synthetic",
        "config": {"workload": w["label"], "sampler": head, "global_batch": len(w["vals"]),
                   "parallelism": "single-gpu"},
        "row_updates_per_sec": sum(w["dims"]) / sec,
        "phase_ms": phase_ms,
        "roofline": roof,
    }
    if w["side"][0] is not None and st.get("univariate"):
        # the row prior update is dominated by the per-latent feature sweeps
        F = w["side"][0]
        out["univariate_sweep"] = {"features": F.shape[1], "latents": K, "prior_rows_ms": float(ph[nm]),
                                   "us_per_feature_step": float(ph[nm]) * 1e3 / max(F.shape[1], 1),
                                   "note": "one wave per latent walks the features in order; time includes the "
                                           "Normal-Wishart draw, the parallel pre-pass, Uhat and lambda_beta"}
    elif w["side"][0] is not None and w["side"][0].shape[1] <= MAX_FF_SIZE:
        nf = w["side"][0].shape[1]
        solver = res[head]["cg"][-2]
        if solver == 2:
            # (F^T F + lambda_beta I)^{-1} as Q diag(d + lambda_beta)^{-1} Q^T: two n x n FP64
            # products per iteration stream Q and Q^T once each
            qb = 2.0 * nf * nf * 8
            gbs = qb / (ph[nm] * 1e-3) / 1e9
            out["ftf_spectral"] = {"features": nf, "prior_rows_ms": float(ph[nm]), "eigenvector_bytes": qb,
                                   "setup_s_incl_eigendecomposition": res[head]["setup_s"],
                                   "achieved_GBs_lower_bound": gbs, "frac_of_hbm_peak": gbs / PEAK_HBM_GBS,
                                   "note": "time includes the two products, the beta noise draw, F^T Y, Uhat "
                                           "and lambda_beta updates and the Normal-Wishart draw"}
        else:
            # the row prior update is dominated by the Cholesky of F^T F + lambda_beta I
            tf = (nf ** 3 / 3.0) / (ph[nm] * 1e-3) / 1e12
            out["ftf_cholesky"] = {"features": nf, "prior_rows_ms": float(ph[nm]), "factor_flops": nf ** 3 / 3.0,
                                   "achieved_TFLOPs_lower_bound": tf, "frac_of_fp64_peak": tf / PEAK_FP64_TFLOPS,
                                   "note": "time includes F^T F + lambda I, the two K-column solves, the beta "
                                           "noise draw, Uhat and lambda_beta updates"}
    elif w["side"][0] is not None:
        # the row prior update is dominated by the block CG for the link matrix
        iters = res[head]["cg"][0]
        cgb = cg_bytes_per_iteration(w) * max(iters, 1)
        gbs = cgb / (ph[nm] * 1e-3) / 1e9
        out["block_cg"] = {"iters_per_solve": iters, "prior_rows_ms": float(ph[nm]),
                           "algorithmic_bytes_per_solve": cgb, "achieved_GBs_lower_bound": gbs,
                           "frac_of_hbm_peak": gbs / PEAK_HBM_GBS,
                           "note": "time includes the beta noise draw, Uhat and lambda_beta updates"}
    if tuning is not None:
        out["tuning"] = tuning
    if head.startswith("batched"):
        st_ = dict(res[head]["cg"][-1])
        out["launch_shapes"] = st_.pop("launch_shapes")
        out["refinement"] = dict(st_, threshold="estimated condition number of the Jacobi-scaled precision > 500, "
                                                "rows of >= max(K, 128) nonzeros")
    if "plain" in res and head != "plain":
        out["plain_iters_per_sec"] = 1.0 / res["plain"]["sec"]
        out["speedup_vs_plain"] = res["plain"]["sec"] / sec
    if not args.no_chain:
        key = "auc" if st["precision"] == "probit" else "rmse"
        out[key] = {k: v.get("rmse") for k, v in res.items()}
        if "plain" in res and head != "plain":
            out[key]["rel_diff_vs_plain"] = res[head]["rmse"] / res["plain"]["rmse"] - 1.0
        if gate:
            out[key]["fp64"] = fp64["rmse"]
            out[key]["rel_diff_vs_fp64"] = {k: v["rmse"] / fp64["rmse"] - 1.0 for k, v in res.items()}
            out[key]["gate"] = "posterior-mean test RMSE of the %d-iteration chain within +0.5%% of the FP64 " \
                               "oracle's on the same data and seed" % (st["burnin"] + st["nsamples"])
            out[key]["gate_met"] = bool(res[head]["rmse"] <= 1.005 * fp64["rmse"])
            out["fp64_error_after_2_iterations"] = {k: v["fp64_err"] for k, v in res.items()}
    if not args.no_cpu and args.workload in ("c1", "c2", "c3", "c4"):
        # a bounded sample of the same workload on the host cores (10-30 s)
        iters = {"c1": st["burnin"] + st["nsamples"], "c2": args.cpu_iters, "c3": 5, "c4": 2}[args.workload]
        out["cpu_baseline"] = cpu_baseline(w, iters, args.seed, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
