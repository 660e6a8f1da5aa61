#!/usr/bin/env python3
"""Benchmark: recovered problems/s for CGIHT on 'gen' 2048 x 8192 (k = 205,
Gaussian nonzeros, FP32), one MI355X.

A step is one pass of the hot path over one batch of synthetic input: one
persistent launch that solves `--batch` resident instances (one workgroup per
instance) from x = 0 to the stopping rule.  value = recovered instances
(relative l_inf error < 1e-3, GAGA/CGIHTcode/shared/tests_delta_fixed.m:17-23)
per second of timed wall clock, inputs already resident in HBM.

Also reported (same JSON line):
  roofline      the solve kernel's algorithmic operator bytes per launch /
                its average launch duration (HIP events), against 8 TB/s
  matvec        the dense A^T r product alone on one instance (whole grid)
  forms         plain (one instance per launch, host-stepped iterations) and
                single (one instance per cooperative launch) problems/s
  cpu_baseline  the FP64 CPU oracle (a C restatement of the reference's
                algorithm) on a bounded sample of the same instances
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20260101
M, N, K = 2048, 8192, 205
TOL = 1e-5
SUCCESS = 1e-3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_baseline(sample):
    """FP64 oracle on `sample` instances of the workload (bounded CPU work)."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ok, t_total = 0, 0.0
    for inst in range(sample):
        P = O.make_problem("gen", M, N, SEED, inst)
        x = O.sparse_vector(N, K, "gaussian", SEED, inst)
        y = P.apply_A(x).astype(np.float32)
        t0 = time.perf_counter()
        xo, res, _, _ = O.solve("CGIHT", P, y, K, tol=TOL, precision=64)
        t_total += time.perf_counter() - t0
        err = np.max(np.abs(xo - x)) / np.max(np.abs(x))
        ok += int(err < SUCCESS)
    return dict(value=ok / t_total, unit="problems/s", cores=threads, kind="port",
                sample="%d CGIHT gen %dx%d k=%d instances, FP64 oracle, %.1f s" % (sample, M, N, K,
                                                                                 t_total))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256, help="instances per step")
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--team", type=int, default=1)
    ap.add_argument("--plain-instances", type=int, default=8)
    ap.add_argument("--cpu-sample", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(instances shard across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import gaga_amd as ga
    ga.load()
    opts = ga.options(tol=TOL, seed=SEED, vecDistribution="gaussian", teamSize=args.team,
                      blockSize=args.block)
    specs = [dict(ens="gen", k=K, m=M, n=N, instance=i) for i in range(args.batch)]
    log("generating %d instances of %dx%d ..." % (args.batch, M, N))
    B = ga.Batch("CGIHT", specs, opts)
    for _ in range(args.warmup):
        B.solve()
    sync()
    t0 = time.perf_counter()
    dev_ms, ok_total, bytes_total = [], 0, 0.0
    for step in range(args.steps):
        dev_ms.append(B.solve())
        r = B.results()
        ok_total += int((r["norms"][:, 2] < SUCCESS).sum())
        bytes_total += B.bytes()
        log("step %d: %.2f ms, %d / %d recovered" % (step, dev_ms[-1], int((r["norms"][:, 2] < SUCCESS).sum()), args.batch))
    sync()
    wall = time.perf_counter() - t0
    r = B.results()
    info = B.info()
    value = ok_total / wall
    mean_ms = float(np.mean(dev_ms))
    bytes_per_launch = bytes_total / args.steps
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    iters = r["iters"]
    linf = r["norms"][:, 2]
    B.close()

    # dense A^T r alone (whole grid on one resident instance)
    pl = ga.Plan("CGIHT", "gen", K, M, N, 0, ga.options(tol=TOL, seed=SEED, instance=0,
                                                       vecDistribution="gaussian"))
    mv_ms, mv_bytes = pl.apply_timed(1, 50)
    # single-instance forms: one instance per launch (device loop) and plain
    # (one outer iteration per launch, host-side stop test)
    forms = {}
    for mode in ("device", "plain"):
        ok, tsum = 0, 0.0
        for inst in range(args.plain_instances):
            p = ga.Plan("CGIHT", "gen", K, M, N, 0, ga.options(
                tol=TOL, seed=SEED, instance=inst, vecDistribution="gaussian", mode=mode))
            p.solve()
            sync()
            t1 = time.perf_counter()
            p.solve()
            sync()
            tsum += time.perf_counter() - t1
            ok += int(p.results()["norms"][2] < SUCCESS)
            p.close()
        forms["single_" + mode] = dict(problems_per_s=ok / tsum, instances=args.plain_instances)
    pl.close()

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_sample)

    out = {
        "metric": "recovered problems/s (rel. linf err < 1e-3), CGIHT gen 2048x8192 k=205, 1 GPU",
        "value": value,
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox4x32-10 instances, A ~ N(0,1/m), x k-sparse N(0,1), y = A x",
        "config": {
            "workload": "CGIHT gen m=2048 n=8192 k=205 gaussian, tol 1e-5, maxiter 5000",
            "batch_instances_per_step": args.batch,
            "form": "batched persistent kernel, %d workgroup(s) x %d threads per instance"
                    % (info["team_size"], info["block"]),
            "teams": info["teams"],
            "success_rate": float((linf < SUCCESS).mean()),
            "iterations_mean": float(iters.mean()),
            "rel_linf_err_median": float(np.median(linf)),
            "forms": forms,
            "plain_to_batched_speedup": value / forms["single_plain"]["problems_per_s"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "solve_batch_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": mean_ms,
        },
        "matvec": {
            "kernel": "apply_coop_kernel (gen A^T r, one instance, whole grid)",
            "achieved": mv_bytes / (mv_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": mv_bytes / (mv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bytes": mv_bytes,
            "avg_ms": mv_ms,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
