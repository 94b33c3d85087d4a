#!/usr/bin/env python
"""Benchmark: dense Rank DP states/s on one MI355X for the single large dense instance, with the roofline of the
dominant kernel and the reference's CPU kernels timed on this host beside it.

Workload (one "step" = one exact distance of the instance: every observable's plan, host planning excluded):
banded random detector error model, seed 2000: 600 detectors, 6000 fault columns each flipping 2-6 detectors inside
a 25-row sliding window, 4 logical rows of density 0.05, integer weights uniform in 1..3, uint8 saturating costs.
The natural column order has rank width 25 and 2^26-entry tables (64 MiB).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]

Prints one JSON line."""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s


def workload():
    from tropidist_amd import problem, trellis
    p = problem.banded(R=600, N=6000, kmin=2, kmax=6, W=25, nobs=4, rho=0.05, weights=(1, 3), seed=2000)
    order = list(range(p.ncols))
    plans = [trellis.build_plan(p.h, p.l, 1 << i, order, p.nrows, weights=p.w) for i in range(p.nobs)]
    return p, [pl for pl in plans if pl.feasible]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--roofline-launches", type=int, default=60)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the sharding design is written up, "
              "not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from tropidist_amd import _lib, kernel
    torch = _lib.torch()
    L = _lib.lib()

    t0 = time.time()
    p, plans = workload()
    plan_seconds = time.time() - t0
    peak = max(pl.peak_bits for pl in plans)
    total_work = sum(pl.work() for pl in plans)
    encoded = [pl.encode() for pl in plans]
    T0 = _lib.device_empty(1 << peak)
    T1 = _lib.device_empty(1 << peak)
    stream = torch.cuda.current_stream()
    sh = _lib.stream_handle(stream)

    def step():
        vals = []
        for pl, ops in zip(plans, encoded):
            v = ctypes.c_int(-1)
            _lib.check(L.tdp_run_plan(ops.ctypes.data_as(ctypes.c_void_p), len(pl.ops), pl.peak_bits,
                                      _lib.dptr(T0), _lib.dptr(T1), ctypes.byref(v), sh), "tdp_run_plan")
            vals.append(kernel.final_value(v.value))
        return vals

    for _ in range(args.warmup):
        values = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        values = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    states_per_s = args.steps * total_work / elapsed

    # Roofline of the dominant kernel (the streaming butterfly update at the peak table size): average launch
    # duration from HIP events on the stream the kernel runs on; algorithmic bytes = read + write of every entry.
    n = 1 << peak
    masks = [op[1] for op in plans[0].ops if op[0] == "update" and op[1] < n and op[1] >= (n >> 1)]
    masks = (masks * (args.roofline_launches // max(1, len(masks)) + 1))[:args.roofline_launches]
    _lib.check(L.tdp_fill(_lib.dptr(T0), n, 7, sh), "fill")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for m in masks:
        _lib.check(L.tdp_update_w(_lib.dptr(T0), n, m, 2, sh), "update")
    ev1.record(stream)
    torch.cuda.synchronize()
    avg_launch = ev0.elapsed_time(ev1) / 1e3 / len(masks)
    alg_bytes = 2 * n
    achieved = alg_bytes / avg_launch

    out = {
        "metric": "Rank-DP states/s (dense trellis, sum of Plan.work / s)",
        "value": states_per_s,
        "unit": "states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded banded detector error model)",
        "config": {"workload": "single large dense Rank DP: banded R=600 N=6000 k=2-6 W=25, 4 logical rows "
                               "rho=0.05, weights 1-3, uint8 saturating, natural order",
                   "columns": p.ncols, "observables": len(plans), "peak_bits": peak, "work_per_step": total_work,
                   "distances": values, "distance": min(values), "host_planning_seconds": round(plan_seconds, 3),
                   "executor": "per-op streaming kernels"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_update_hi", "launch_seconds": avg_launch, "alg_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        import oracle
        if oracle.have_ref():
            R = oracle.rlib("viterbi")
            threads = R.dp_num_threads()
            c0 = time.perf_counter()
            cv = oracle.run_plan_c(plans[0], "ref")
            cpu_s = time.perf_counter() - c0
            out["cpu_baseline"] = {"value": plans[0].work() / cpu_s, "unit": "states/s", "cores": threads,
                                   "kind": "reference",
                                   "sample": "observable 0 of the same instance (work %d states), the reference's "
                                             "viterbi.c (OpenMP) driven op by op" % plans[0].work(),
                                   "seconds": cpu_s, "distance": cv, "matches_gpu": cv == values[0]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
