#!/usr/bin/env python3
"""k-Shape benchmark on one MI355X: n=100000 series of length m=512, k=16
clusters (FFT length 1024), the large single-job workload with the precision
path (FP32 SBD, split-BF16 MFMA for S).

A step is one full k-Shape iteration (shape extraction for all k clusters,
then the fused SBD assignment of all n series) on data resident in HBM.
`value` = SBD pairs per second over full iterations (n*k per iteration).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "times-c_amd"))

METRIC = "SBD pairs/s + k-Shape iter/s at n=100k,m=512,k=16; p50/p99 job latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes per k_assign_prune launch at the default workload (n=100000,
# m=512, k=16), from rocprofv3 FETCH_SIZE (x2) and WRITE_SIZE (x1) counter
# runs calibrated on a plain 256 MiB copy: profiles/kshape_iteration_n100k_m512_k16.md
ASSIGN_TRAFFIC_BYTES = {(100000, 512, 16): (444021 + 9992) * 1024}


def fft_size(m):
    return 1 << (2 * m - 1).bit_length()


def sbd_assign_flops(n, m, k, evals=None):
    """FP32 flops of one assignment, FFTW convention (5 F log2 F per complex
    transform of length F; two real series share one transform).

    Full search: per series a half forward transform and its norm, per
    (series, centroid) a half inverse transform, the conj-multiply of F/2+1
    bins and the 2m-1 lag max.  Pruned search inside a fit (`evals` inverse
    transforms of series pairs were run; the series' half spectra and norms
    are transformed once per fit, and the bounds of every (series, centroid)
    come from a separate product kernel before the assignment): per pair the
    packed spectrum from two half spectra (2 F), per evaluation the
    conj-multiply (6 F), the inverse transform and both lag maxima."""
    F = fft_size(m)
    t = 5.0 * F * math.log2(F)
    if evals is None:
        return n * (t / 2 + 2 * m) + n * k * (t / 2 + 6 * (F // 2 + 1) + 2 * m - 1)
    pairs = (n + 1) // 2
    return pairs * 2 * F + evals * (t + 6 * F + 4 * m)


def assign_bytes(n, m, k):
    """Algorithmic HBM bytes of one pruned assignment inside a fit: the
    series' half spectra (F/2+1 complex FP32 each), norms and k bounds, the
    old labels read and labels, distances and shifts written."""
    F = fft_size(m)
    return n * (F // 2 + 1) * 8 + n * 4 + n * k * 4 + 4 * n * 4


def iteration_bytes(n, m, k):
    """Compulsory HBM bytes of one iteration: X read by refine and by assign,
    label/dist/shift, centroids and their spectra."""
    F = fft_size(m)
    return 2 * n * m * 4 + 3 * n * 4 + k * m * 4 + k * (F // 2 + 1) * 8


def cpu_baseline(X, k, sample, seed):
    """FP64 oracle (the C restatement of the reference's CPU path) on a bounded
    sample of the same workload, all host cores: one refine + one assign."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    import timesc

    Xs = oracle_py.zscore(X[:sample].astype(np.float64))
    lab = timesc.init_labels(seed, sample, k)
    C0 = np.zeros((k, X.shape[1]))
    threads = os.cpu_count() or 1
    # untimed first iteration (zero centroids, no alignment), then one timed
    # iteration with alignment, like the GPU's timed steps
    C, _, _ = oracle_py.refine(Xs, lab, C0, seed=seed, it=0, nthreads=threads)
    lab, _, _ = oracle_py.assign(Xs, C, nthreads=threads)
    t0 = time.perf_counter()
    C, _, _ = oracle_py.refine(Xs, lab, C, seed=seed, it=1, nthreads=threads)
    t1 = time.perf_counter()
    oracle_py.assign(Xs, C, nthreads=threads)
    t2 = time.perf_counter()
    pairs = sample * k
    return {
        "value": pairs / (t2 - t0),
        "unit": "SBD pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"iteration 2 (refine {t1 - t0:.1f} s + assign {t2 - t1:.1f} s) on the first "
                  f"{sample} series (m={X.shape[1]}, k={k}), FP64 oracle, OpenMP",
        "assign_pairs_per_s": pairs / (t2 - t1),
        "iter_per_s": 1.0 / (t2 - t0),
    }


def batch_workload(timesc, njobs):
    """Uneven independent fits: job j has n ~ U{100..5000}, m = round(exp(U[ln 64,
    ln 2048])), k ~ U{2..32}, the benchmark generator at sigma 0.3 with seed
    1000 + j, random initial labels, at most 10 iterations."""
    jobs = []
    for j in range(njobs):
        rng = np.random.default_rng(1000 + j)
        n = int(rng.integers(100, 5001))
        m = int(round(math.exp(rng.uniform(math.log(64), math.log(2048)))))
        k = int(rng.integers(2, 33))
        X, _ = timesc.generate(1000 + j, n, m, k, 0.3)
        jobs.append((X, k, 1000 + j))
    return jobs


def run_batch(timesc, ctx, jobs, workers):
    """Runs the uneven batch twice (the first run sizes every worker's
    workspace) and reports the second: per-job latency from batch start to
    job completion, per-job run time, jobs/s."""
    ds = [ctx.to_device(X) for X, _, _ in jobs]
    specs = [{"d": d, "k": k, "seed": seed, "znorm_ddof": 0, "max_iter": 10}
             for d, (_, k, seed) in zip(ds, jobs)]
    try:
        with timesc.Batch(0, workers=workers) as b:
            b.run(specs, order=timesc.ORDER_LONGEST_FIRST)
            for d, (X, _, _) in zip(ds, jobs):  # the first run normalized in place
                ctx.h2d(d, X)
            t0 = time.perf_counter()
            res = b.run(specs, order=timesc.ORDER_LONGEST_FIRST)
            wall = time.perf_counter() - t0
    finally:
        for d in ds:
            d.free()
    lat = np.array([r["ms_latency"] for r in res])
    run = np.array([r["ms_run"] for r in res])
    iters = np.array([r["stats"]["iterations"] for r in res])
    return {
        "jobs": len(jobs), "workers": workers, "makespan_ms": wall * 1e3,
        "jobs_per_s": len(jobs) / wall,
        "p50_latency_ms": float(np.percentile(lat, 50)),
        "p99_latency_ms": float(np.percentile(lat, 99)),
        "p50_run_ms": float(np.percentile(run, 50)), "p99_run_ms": float(np.percentile(run, 99)),
        "mean_iterations": float(iters.mean()),
        "workload": "n~U{100..5000}, m=round(exp U[ln64,ln2048]), k~U{2..32}, sigma 0.3, "
                    "seeds 1000+j, <=10 iterations, z-norm included, all jobs submitted at once",
    }


def serve_workload(timesc, nreq, qps, seed=7):
    """Poisson arrivals at `qps`: 30% fits (sizes of the uneven batch), 70%
    assignment queries of 1..1024 series against one of 8 resident models
    (m = round(exp(U[ln 64, ln 1024])), k ~ U{2..64})."""
    rng = np.random.default_rng(seed)
    models = []
    for j in range(8):
        m = int(round(math.exp(rng.uniform(math.log(64), math.log(1024)))))
        k = int(rng.integers(2, 65))
        C, _ = timesc.generate(2000 + j, k, m, k, 0.3)
        C = (C - C.mean(1, keepdims=True)) / C.std(1, keepdims=True)
        models.append(np.ascontiguousarray(C, np.float32))
    reqs = []
    t = 0.0
    for i in range(nreq):
        t += rng.exponential(1.0 / qps)
        if rng.random() < 0.3:
            n = int(rng.integers(100, 5001))
            m = int(round(math.exp(rng.uniform(math.log(64), math.log(2048)))))
            k = int(rng.integers(2, 33))
            reqs.append((t, "fit", timesc.generate(3000 + i, n, m, k, 0.3)[0], k))
        else:
            j = int(rng.integers(0, len(models)))
            n = int(rng.integers(1, 1025))
            reqs.append((t, "assign", timesc.generate(3000 + i, n, models[j].shape[1], 4, 0.3)[0],
                         j))
    return models, reqs


def run_serve(timesc, models, reqs, qps, fit_workers, query_workers):
    """Replays the request stream twice on one server: once submitted at
    once (sizes every worker's workspace), then at its arrival times.
    Latency = submit -> result, per request kind."""
    def submit(srv, mids, r):
        _, kind, X, a = r
        return srv.submit_fit(X, a) if kind == "fit" else srv.submit_assign(mids[a], X)

    with timesc.Server(0, fit_workers=fit_workers, query_workers=query_workers) as srv:
        mids = [srv.add_model(C) for C in models]
        for h in [submit(srv, mids, r) for r in reqs]:
            srv.wait(h)
        hs = []
        t0 = time.perf_counter()
        for r in reqs:
            delay = t0 + r[0] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            hs.append((r[1], submit(srv, mids, r)))
        res = [(kind, srv.wait(h)) for kind, h in hs]
        wall = time.perf_counter() - t0

    def pct(kind):
        lat = np.array([r["ms_latency"] for k, r in res if kind in (None, k)])
        return {"count": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99))}

    return {
        "requests": len(reqs), "qps_offered": qps, "qps_served": len(reqs) / wall,
        "fit_workers": fit_workers, "query_workers": query_workers,
        "fit": pct("fit"), "assign": pct("assign"), "all": pct(None),
        "workload": "Poisson arrivals; 30% fits (n~U{100..5000}, m=round(exp U[ln64,ln2048]), "
                    "k~U{2..32}, <=10 iterations), 70% assignment queries (1..1024 series) "
                    "against 8 resident models (m<=1024, k~U{2..64}); z-norm included",
    }


def copy_calibration(mib):
    """A plain copy of `mib` MiB (read once, written once), three launches."""
    import torch

    x = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": x.numel() * 4, "launches": 3}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--m", type=int, default=512)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=100000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-workers", type=int, default=8)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-qps", type=float, default=100.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times "
                         "(the reference kernel for the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; run with --gpus 1", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    import timesc

    n, m, k = args.n, args.m, args.k
    print(f"# generating n={n} m={m} k={k}", file=sys.stderr, flush=True)
    X, y = timesc.generate(args.seed, n, m, k, 0.5)
    ctx = timesc.Context(0)
    dx = ctx.to_device(X)
    ctx.znorm(dx, ddof=0)
    ctx.fit_begin(dx, k, max_iter=1 << 30, seed=args.seed)
    for _ in range(args.warmup):
        ctx.fit_step()
    ctx.sync()
    s0 = ctx.fit_query()
    print(f"# warmup done ({args.warmup} iterations)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.fit_step()
    ctx.sync()
    t1 = time.perf_counter()
    s1 = ctx.fit_query()
    res = ctx.fit_end(n, k, m)
    dt = t1 - t0
    ms_step = dt / args.steps * 1e3
    ms_assign = (s1["ms_assign"] - s0["ms_assign"]) / args.steps
    ms_bounds = (s1["ms_assign_bounds"] - s0["ms_assign_bounds"]) / args.steps
    ms_refine = (s1["ms_refine"] - s0["ms_refine"]) / args.steps
    evals = (s1["sbd_evals"] - s0["sbd_evals"]) / args.steps
    flops = sbd_assign_flops(n, m, k, evals)
    # the pruned kernel's own launch: the assignment minus the bound product
    ms_kernel = ms_assign - ms_bounds
    achieved = flops / (ms_kernel * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": n * k * args.steps / dt,
        "unit": "SBD pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"k-Shape iteration, n={n} m={m} k={k} F={fft_size(m)}, sigma=0.5, "
                        "FP32 SBD, split-BF16 MFMA S",
            "n": n, "m": m, "k": k, "fft": fft_size(m), "s_precision": "split-bf16",
        },
        "iter_per_s": args.steps / dt,
        "assign_ms": ms_assign,
        "assign_bounds_ms": ms_bounds,
        "refine_ms": ms_refine,
        "assign_pairs_per_s": n * k / (ms_assign * 1e-3),
        "hbm_fraction_of_iteration": iteration_bytes(n, m, k) / (PEAK_HBM_GBS * 1e9) / (ms_step * 1e-3),
        "power_steps_max": s1["power_steps_max"],
        "assign_eval_fraction": evals / (((n + 1) // 2) * k),
        "assign_full_search_equiv_tflops": sbd_assign_flops(n, m, k) / (ms_assign * 1e-3) / 1e12,
        "p50_job_ms": None,
        "p99_job_ms": None,
        "roofline": {
            "bound": "mfma",
            "kernel": "k_assign_prune (fused FFT SBD assignment, FP32 VALU)",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": ASSIGN_TRAFFIC_BYTES.get((n, m, k)),
            "algorithmic_bytes": assign_bytes(n, m, k),
            "note": "compute-bound FP32 VALU (FFT butterflies, no matrix product); the FP32 "
                    "peak counts packed FP32, which this build disables (78.6 TFLOP/s unpacked)",
        },
        "cpu_baseline": None,
    }
    dx.free()
    if not args.no_batch and args.batch_jobs > 0:
        print(f"# batch of {args.batch_jobs} uneven fits", file=sys.stderr, flush=True)
        b = run_batch(timesc, ctx, batch_workload(timesc, args.batch_jobs), args.batch_workers)
        out["batch"] = b
        out["p50_job_ms"] = b["p50_latency_ms"]
        out["p99_job_ms"] = b["p99_latency_ms"]
    if not args.no_serve and args.serve_requests > 0:
        print(f"# serving {args.serve_requests} requests at {args.serve_qps} qps", file=sys.stderr,
              flush=True)
        models, reqs = serve_workload(timesc, args.serve_requests, args.serve_qps)
        out["serve"] = run_serve(timesc, models, reqs, args.serve_qps, 8, 4)
    ctx.close()
    if not args.no_cpu_baseline:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(X, k, min(args.cpu_sample, n), args.seed)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
