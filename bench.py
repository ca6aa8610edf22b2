#!/usr/bin/env python3
"""Benchmark: Gibbs iterations/s of the BPMF sampler on one MI355X.

Workload: the 138,000 x 27,000 recommender-scale BPMF with 20M observed cells
(18M train, 2M test), K = 32, adaptive noise precision (macau_amd.synthetic.c2).
A step is one full Gibbs iteration (both half-sweeps, both Normal-Wishart
draws, the noise update and the test evaluation) with all inputs resident in
HBM.  The headline value is the batched mixed-precision sampler; the plain
FP32 sampler is timed beside it on the same data, and the posterior-mean test
RMSE of both full chains (burnin 50 + 100 samples) is reported.

The dominant kernel group is the per-row latent sampler (both half-sweeps);
its duration is measured with HIP events on the model's stream
(macau_profile) and priced against the MFMA/vector peak of the arithmetic it
uses: algorithmic FLOPs per half-sweep = nnz (K^2 + 3K) + rows (K^3/3 + 4K^2).

cpu_baseline: the FP64 C/OpenMP oracle timed on this host for a bounded
sample of the same workload (25 full Gibbs iterations, about 10 s on 16
cores).
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
PEAK_TFLOPS = {"plain": 157.3, "batched": 2500.0}  # FP32 vector/MFMA; BF16 dense MFMA
UNIT_OF = {"plain": "f32", "batched": "bf16x3-f32acc"}


def build_model(train, test, shape, settings, sampler, seed, graph=True):
    L = _lib.lib()
    K = settings["num_latent"]
    if settings["precision"] == "adaptive":
        h = L.macau_make_adaptive(2, K, 1.0, 10.0)
    else:
        h = L.macau_make_fixed(2, K, float(settings["precision"]))
    if not h:
        raise RuntimeError(_lib.last_error())
    r, c, v = train
    idx = np.asfortranarray(np.stack([r, c], 1).astype(np.int32))
    dims = np.array(shape, dtype=np.int32)
    check(L.macau_set_relation_data(h, ptr(idx), 2, ptr(np.ascontiguousarray(v, dtype=np.float64)), len(v),
                                    ptr(dims)))
    rt, ct, vt = test
    tidx = np.asfortranarray(np.stack([rt, ct], 1).astype(np.int32))
    check(L.macau_set_relation_data_test(h, ptr(tidx), 2, ptr(np.ascontiguousarray(vt, dtype=np.float64)),
                                         len(vt), ptr(dims)))
    check(L.macau_add_bpmf_prior(h))
    check(L.macau_add_bpmf_prior(h))
    check(L.macau_set_samples(h, settings["burnin"], settings["nsamples"]))
    check(L.macau_set_seed(h, seed))
    check(L.macau_set_sampler(h, sampler))
    check(L.macau_set_graph(h, int(graph)))
    check(L.macau_set_verbose(h, 0))
    check(L.macau_init(h))
    return h


def sampler_flops(train, shape, K):
    r, c, _ = train
    nnz = len(r)
    per_nnz = K * K + 3 * K
    per_row = K ** 3 / 3.0 + 4 * K * K
    return 2 * nnz * per_nnz + (shape[0] + shape[1]) * per_row


def time_sampler(name, train, test, shape, settings, steps, warmup, seed, log):
    L = _lib.lib()
    kind = _lib.SAMPLER_BATCHED if name == "batched" else _lib.SAMPLER_PLAIN_FP32
    h = build_model(train, test, shape, settings, kind, seed)
    try:
        check(L.macau_iterate(h, warmup))
        check(L.macau_synchronize(h))
        t0 = time.perf_counter()
        check(L.macau_iterate(h, steps))
        check(L.macau_synchronize(h))
        t1 = time.perf_counter()
        nprof = max(3, min(steps, 10))
        phase = np.zeros(2 + 3)
        check(L.macau_profile(h, nprof, ptr(phase)))
        log("%s: %d steps in %.3f s -> %.1f it/s; phases(ms) %s" %
            (name, steps, t1 - t0, steps / (t1 - t0), np.round(phase, 3).tolist()))
    finally:
        L.macau_free(h)
    return (t1 - t0) / steps, phase


def full_chain_rmse(name, train, test, shape, settings, seed):
    L = _lib.lib()
    kind = _lib.SAMPLER_BATCHED if name == "batched" else _lib.SAMPLER_PLAIN_FP32
    h = build_model(train, test, shape, settings, kind, seed)
    try:
        n = settings["burnin"] + settings["nsamples"]
        t0 = time.perf_counter()
        check(L.macau_iterate(h, n))
        check(L.macau_synchronize(h))
        t1 = time.perf_counter()
        return L.macau_get_rmse_avg(h), n / (t1 - t0)
    finally:
        L.macau_free(h)


def cpu_baseline(train, test, shape, settings, iters, seed, log):
    import oracle as O  # the FP64 oracle, as the CPU baseline
    threads = os.cpu_count() or 1
    O.set_threads(threads)
    r, c, v = train
    modes, mean, nnz = O.matrix_rows(r, c, v, shape[0], shape[1])
    ti, tv = O.matrix_test(test[0], test[1], test[2], shape[0], shape[1])
    t0 = time.perf_counter()
    O.run(modes, mean, nnz, list(shape), settings["num_latent"], ti, tv, adaptive=True,
          burnin=settings["burnin"], nsamples=settings["nsamples"], seed=seed, max_iters=iters)
    dt = time.perf_counter() - t0
    log("cpu oracle: %d iterations in %.2f s on %d threads" % (iters, dt, threads))
    return {"value": iters / dt, "unit": "iters/s", "cores": O.get_threads(), "kind": "port",
            "sample": "%d full Gibbs iterations of the same 138000x27000 / 18M-train K=32 workload, FP64, "
                      "OpenMP dynamic,2 rows (C restatement of the reference)" % iters}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chain", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=25)
    ap.add_argument("--sampler", choices=["batched", "plain"], default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (one GPU per box)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    train, test, shape, settings = synthetic.c2(args.seed)
    K = settings["num_latent"]
    log("generated C2 in %.1f s: %d train, %d test" % (time.perf_counter() - t0, len(train[2]), len(test[2])))

    head = args.sampler or "batched"
    res = {}
    names = [head] + ([] if (args.no_plain or head == "plain") else ["plain"])
    for name in names:
        sec, phase = time_sampler(name, train, test, shape, settings, args.steps, args.warmup, args.seed, log)
        res[name] = {"sec": sec, "phase": phase}
        if not args.no_chain:
            rmse, ips = full_chain_rmse(name, train, test, shape, settings, args.seed)
            res[name]["rmse"] = rmse
            log("%s full chain (%d iterations): posterior-mean RMSE %.6f" %
                (name, settings["burnin"] + settings["nsamples"], rmse))

    flops = sampler_flops(train, shape, K)
    ph = res[head]["phase"]
    sampler_ms = ph[0] + ph[1]
    achieved = flops / (sampler_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS[head], "traffic": load_traffic(head),
            "kernel": "per-row latent sampler, both half-sweeps (%s)" % head,
            "flops_per_iteration": flops, "ms_per_iteration": sampler_ms}
    sec = res[head]["sec"]
    out = {
        "metric": METRIC,
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
        "config": {"workload": "BPMF 138000x27000, 20M observed (18M train / 2M test), Zipf-1.1 degrees, "
                               "K=32, adaptive noise", "sampler": head, "global_batch": len(train[2]),
                   "parallelism": "single-gpu"},
        "row_updates_per_sec": (shape[0] + shape[1]) / sec,
        "phase_ms": {"sample_rows": ph[0], "sample_cols": ph[1], "priors": ph[2], "noise": ph[3], "eval": ph[4]},
        "roofline": roof,
    }
    if "plain" in res and head != "plain":
        out["plain_iters_per_sec"] = 1.0 / res["plain"]["sec"]
        out["speedup_vs_plain"] = res["plain"]["sec"] / sec
    if not args.no_chain:
        out["rmse"] = {k: v.get("rmse") for k, v in res.items()}
        if "plain" in res and head != "plain":
            out["rmse"]["rel_diff_vs_plain"] = res[head]["rmse"] / res["plain"]["rmse"] - 1.0
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(train, test, shape, settings, args.cpu_iters, args.seed, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
