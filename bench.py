#!/usr/bin/env python
"""Benchmark of the exact distance evaluators on one MI355X.  Prints one JSON line.

value = dense Rank DP states/s (sum of Plan.work per second) on the single large dense instance: banded random
detector error model, seed 2000: 600 detectors, 6000 fault columns each flipping 2-6 detectors inside a 25-row
sliding window, 4 logical rows of density 0.05, integer costs uniform in 1..3, uint8 saturating tables; the natural
column order has rank width 25 and 2^26-entry tables.  One step = the exact distance of the instance: the four
observables' fused programs run together, one pass of each per round (host planning and pass compilation are
timed separately).  roofline: the widest passes of observable 0 (the dominant kernel), each replayed by HIP events,
averaged; roofline.dram: the same for the W = 27 instance, whose 2^28-entry tables do not fit the MALL.

Beside it (each can be skipped):
  batched   512 banded scans (200-1200 detectors, 5-15 faults per detector, width 10-22, half with costs 1-3):
            certificates (distance + checked witness) per second (the median of five passes over the scan), against
            the same plans certified one at a time by the single-plan fused executors ("plain GPU", whole scan);
  sparse    the two-block quasi-cyclic code l = m = 12 (288 columns), meet-in-the-middle sweep at threshold K, and
            a reference-sized code (l = m = 9) against the reference's sparse_dp.c on the host;
  tree      the tree-realization DP of a banded model (R=75 N=300 W=10) on a balanced tree: join pairs per second,
            against the reference's tree_dp.c joins on the host;
  tables    the meet-in-the-middle lookup table of every column subset of size <= 3 of the l = m = 12 code: build
            time per entry and batched lookups per second, against the reference's gmitm.c build on the host;
  serving   the serving mix of 10,000 requests (70% dense width 10-20, 20% width 21-25, 10% sparse MITM with
            K <= 10 on quasi-cyclic codes): capacity with every request submitted at once, then the first 2,000 as a
            Poisson stream at half of it, p50 / p99 latency, against a stratified plain-GPU sample;
  cpu_configs  the reference's kernels on the host: the plain dense DP on one thread (R=48 N=300 W=10), and
            certificates per second of the batched scan and the serving mix on all host cores;
  cpu_baseline  the reference's own viterbi.c (compiled from its sources, OpenMP on this host) on observable 0.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--skip-batched] [--skip-sparse]
                    [--skip-tree] [--skip-tables] [--skip-serving] [--serving-requests 10000] [--serving-max-K 10]
                    [--serving-max-states 4000000] [--serving-poisson-requests 2000] [--batch-instances 512]
                    [--sparse-K 12]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
T_START = time.time()


def _gen_scan(i):
    from tropidist_amd import workloads
    import numpy as np
    rng = np.random.default_rng(4000 + i)
    return workloads.banded_instance(rng, 4000 + i + 1, int(rng.integers(10, 23)), weighted=bool(i % 2))


def _gen_request(req):
    from tropidist_amd import problem, workloads
    import numpy as np
    kind, a = req
    if kind == "sparse":
        return kind, problem.quasi_cyclic(a["l"], a["m"], seed=a["seed"]), a["K"]
    rng = np.random.default_rng(a["seed"])
    p = workloads.banded_instance(rng, a["seed"], a["width"], a["weighted"])
    p.band_width = a["width"]                   # the request's class (narrow below 21, wide from 21)
    return kind, p, None


def generate(args):
    """Synthetic inputs of every section, made in worker processes before the GPU is touched."""
    import multiprocessing as mp
    from tropidist_amd import workloads
    out = {}
    reqs = workloads.serving_mix(args.serving_requests, K=(4, args.serving_max_K))
    if args.gen_workers <= 0:                   # in process (profiled runs: no child processes)
        if not args.skip_batched:
            out["scan"] = [_gen_scan(i) for i in range(args.batch_instances)]
        if not args.skip_serving:
            out["requests"] = [_gen_request(r) for r in reqs]
        return out
    ctx = mp.get_context("spawn")
    with ctx.Pool(min(args.gen_workers, os.cpu_count() or 4)) as pool:
        if not args.skip_batched:
            out["scan"] = pool.map(_gen_scan, range(args.batch_instances), chunksize=8)
        if not args.skip_serving:
            out["requests"] = pool.map(_gen_request, reqs, chunksize=8)
    return out


def _time_launches(torch, stream, launch, count):
    """Average duration of `count` back-to-back launches on `stream`, by HIP events recorded on that stream."""
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    launch()
    ev0.record(stream)
    for _ in range(count):
        launch()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / 1e3 / count


def dense_setup(L, torch, W=25, nobs=None):
    """The headline instance (window W, its first `nobs` observables), its plans and fused programs, and device
    tables (shared with the profiling mode)."""
    from tropidist_amd import _lib, fused, trellis, workloads
    t0 = time.time()
    p = workloads.single_large(W=W)
    order = list(range(p.ncols))
    obs = [1 << i for i in range(p.nobs if nobs is None else nobs)]
    plans = [pl for pl in trellis.build_plans_native(p.h, p.l, obs, order, p.nrows, p.w) if pl.feasible]
    plan_seconds = time.time() - t0
    t0 = time.time()
    progs = [fused.program_for(pl, min_bits=2) for pl in plans]      # the programs fused.run_many runs
    compile_seconds = time.time() - t0
    peak = max(pl.peak_bits for pl in plans)
    T0, T1 = _lib.device_empty(1 << peak), _lib.device_empty(1 << peak)
    return p, plans, progs, peak, T0, T1, plan_seconds, compile_seconds


def dominant_passes(progs):
    """(kernel name, algorithmic bytes, descriptor records) of the dominant kernel: every pass of the first
    program (observable 0) that moves the most bytes (the widest table) and runs the most common kernel among
    those."""
    import collections
    prog = progs[0]
    top = max(prog.pass_bytes(i) for i in range(prog.npasses))
    widest = [i for i in range(prog.npasses) if prog.pass_bytes(i) == top]
    kname = collections.Counter(prog.pass_kernel(i) for i in widest).most_common(1)[0][0]
    idx = [i for i in widest if prog.pass_kernel(i) == kname]
    return kname, top, prog.arr[idx].copy()


def measured_traffic(name, kernel, alg_bytes):
    """HBM bytes per launch from a committed counter summary (profiles/<name>.json: rocprofv3 FETCH_SIZE and
    WRITE_SIZE, scaled by a device copy of known size), when it was measured on the same kernel and algorithmic
    bytes per launch (within 0.1%); None otherwise."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    a = prof.get("alg_bytes_per_launch") or 0
    if kernel not in prof.get("kernel", "") or abs(a - alg_bytes) > 1e-3 * alg_bytes:
        return None
    return prof["traffic_bytes_per_launch"]


def step_roofline(plans, stream, step_seconds):
    """Roofline of the step's dominant kernel, k_pass_round (one launch per round: the next pass of each of the
    four programs, the round's jobs passed by value): one more step with HIP events recorded on the step's stream around every round
    (tdp_run_programs_timed), and each round's algorithmic bytes (every input entry read once, every output entry
    written once).  achieved = all rounds' bytes / all rounds' time; the widest rounds alone beside it."""
    import numpy as np
    from tropidist_amd import fused
    timing = {}
    fused.run_many(plans, stream=stream, timing=timing)
    us, by = timing["round_us"], timing["round_bytes"]
    kernel_s, alg = float(us.sum()) * 1e-6, float(by.sum())
    wide = by >= 0.999 * by.max()
    wide_s = float(us[wide].sum()) * 1e-6
    return {"bound": "hbm", "achieved": alg / kernel_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": alg / kernel_s / HBM_PEAK,
            "traffic": measured_traffic("step_pass_multi_banded_w25", "k_pass_round", alg / len(us)),
            "traffic_source": "profiles/step_pass_multi_banded_w25.json (rocprofv3 FETCH_SIZE / WRITE_SIZE of every "
                              "k_pass_round launch of one step, scaled by a device copy of 2^26 bytes)",
            "kernel": "tdp::k_pass_round", "launches": int(len(us)), "alg_bytes_per_launch": alg / len(us),
            "launch_seconds": kernel_s / len(us), "kernel_seconds_per_step": kernel_s,
            "kernel_share_of_step": kernel_s / step_seconds, "step_alg_bytes": alg,
            "step_frac": alg / step_seconds / HBM_PEAK,
            "widest_rounds": {"launches": int(wide.sum()), "alg_bytes_per_launch": float(by.max()),
                              "launch_seconds": wide_s / max(1, int(wide.sum())),
                              "frac": float(by[wide].sum()) / wide_s / HBM_PEAK},
            "round_bytes_histogram": {str(int(b)): int(c) for b, c in zip(*np.unique(by, return_counts=True))
                                      if c >= 8}}


def bench_dense(args, L, torch):
    import ctypes
    import numpy as np
    from tropidist_amd import _lib, kernel
    p, plans, progs, peak, T0, T1, plan_seconds, compile_seconds = dense_setup(L, torch)
    total_work = sum(pl.work() for pl in plans)
    encoded = [pl.encode() for pl in plans]
    stream = torch.cuda.current_stream()
    sh = _lib.stream_handle(stream)

    from tropidist_amd import fused

    def step():
        # the four observables' fused programs advance together, one pass of each per round (tdp_run_programs);
        # the host schedule of the four is built inside the timed step
        return fused.run_many(plans, stream=stream)

    for _ in range(args.warmup):
        values = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        values = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1

    # the same step captured once as a HIP graph (the schedule built, uploaded and recorded outside the timing) and
    # replayed: the device work of every step again, without the host's per-round launch calls
    graph = fused.ProgramGraph(plans, stream=stream)
    graph.launch(stream)
    graph_values = graph.values(stream)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch(stream)
    torch.cuda.synchronize()
    graph_s = (time.perf_counter() - t3) / args.steps
    graph_values_after = graph.values(stream)
    graph_nodes = graph.nodes
    graph.close()

    # per-op executor on the same step, once, for comparison
    t2 = time.perf_counter()
    per_op_values = []
    for pl, ops in zip(plans, encoded):
        v = ctypes.c_int(-1)
        _lib.check(L.tdp_run_plan(ops.ctypes.data_as(ctypes.c_void_p), len(ops), pl.peak_bits, _lib.dptr(T0),
                                  _lib.dptr(T1), ctypes.byref(v), sh), "tdp_run_plan")
        per_op_values.append(kernel.final_value(v.value))
    per_op_seconds = time.perf_counter() - t2

    # the certificate of the step's distance: the minimising observable's witness at 2^26 entries (fused witness
    # program, every step's input kept in one arena), checked against H and L
    from tropidist_amd import auto
    from tropidist_amd.problem import verify_witness
    best = min(values)
    wobs = values.index(best)             # every observable of this instance has a feasible plan
    assert len(plans) == p.nobs
    torch.cuda.synchronize()
    tw = time.perf_counter()
    wv, wcols = auto.witness_plan(plans[wobs], list(range(p.ncols)), stream=stream)
    torch.cuda.synchronize()
    witness = {"observable": wobs, "value": wv, "columns": len(wcols), "seconds": time.perf_counter() - tw,
               "peak_bits": plans[wobs].peak_bits,
               "witness_ok": wv == best and verify_witness(p, wcols, best, wobs) and p.cost(wcols) == best}
    roof = step_roofline(plans, stream, elapsed / args.steps)
    # beside it: the widest fused passes of observable 0 alone, each replayed on its own input (T0 -> T1) by the
    # single-program kernel, the average launch
    kname, pass_bytes, descs = dominant_passes(progs)
    d = descs[0]
    _lib.check(L.tdp_fill(_lib.dptr(T0), 1 << int(d["b_in"]), 7, sh), "fill")
    per = max(1, args.roofline_launches // len(descs))
    pass_s = sum(_time_launches(torch, stream, lambda k=k: _lib.check(
        L.tdp_run_pass(_lib.dptr(T0), _lib.dptr(T1), descs[k:k + 1].ctypes.data_as(ctypes.c_void_p), sh),
        "tdp_run_pass"), per) for k in range(len(descs))) / len(descs)
    # the per-op streaming update at the peak table, for reference
    n = 1 << peak
    masks = [int(m) for c, m, _ in encoded[0].tolist() if c == 1 and (n >> 1) <= m < n][:1] or [n >> 1]
    upd_s = _time_launches(torch, stream, lambda: _lib.check(L.tdp_update_w(_lib.dptr(T0), n, masks[0], 2, sh),
                                                            "update"), max(1, args.roofline_launches))
    del T0, T1
    dram = dram_pass(args, L, torch, stream, sh)
    return p, plans, values, {
        "value": args.steps * total_work / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
        "config": {"workload": "single large dense Rank DP: banded R=600 N=6000 k=2-6 W=25, 4 logical rows "
                               "rho=0.05, costs 1-3, uint8 saturating, natural order, seed 2000",
                   "columns": p.ncols, "observables": len(plans), "peak_bits": peak, "work_per_step": total_work,
                   "distances": values, "distance": min(values), "host_planning_seconds": round(plan_seconds, 3),
                   "pass_compile_seconds": round(compile_seconds, 3), "witness": witness,
                   "passes_per_step": sum(pr.npasses for pr in progs),
                   "executor": "fused passes, the four programs advanced together one pass each per round "
                               "(tdp_run_programs)",
                   "graph_replay": {"states_per_s": total_work / graph_s, "ms_per_step": 1e3 * graph_s,
                                    "nodes": graph_nodes, "distances": graph_values,
                                    "identical": graph_values == values == graph_values_after},
                   "per_op_executor": {"states_per_s": total_work / per_op_seconds,
                                       "ms_per_step": 1e3 * per_op_seconds, "distances": per_op_values,
                                       "identical": per_op_values == values}},
        "roofline": dict(roof, pass_replay={
            "kernel": kname, "launch_seconds": pass_s, "alg_bytes_per_launch": pass_bytes,
            "achieved": pass_bytes / pass_s / 1e9, "frac": pass_bytes / pass_s / HBM_PEAK,
            "traffic": measured_traffic("pass_d5_banded_w25", kname + "(", pass_bytes),
            "passes": {"count": len(descs), "b_in": int(d["b_in"]), "b_out": int(d["b_out"]),
                       "updates": [int(x["D"]) + int(x["nlow"]) for x in descs[:1]][0],
                       "threads": 1 << int(d["nfree"]), "launches_each": per}},
            per_op_update={"kernel": "tdp::k_update_hi", "table_entries": n, "launch_seconds": upd_s,
                           "achieved": 2 * n / upd_s / 1e9, "frac": 2 * n / upd_s / HBM_PEAK},
            dram=dram),
    }


def dram_pass(args, L, torch, stream, sh, W=27):
    """The same pass kernel on tables too large for the 256 MB MALL: the widest pass of observable 0 of the W = 27
    instance (2^28 -> 2^28 entries, 512 MiB moved per launch), timed by HIP events on its stream."""
    import ctypes
    from tropidist_amd import _lib
    _, _, progs, _, T0, T1, _, _ = dense_setup(L, torch, W=W, nobs=1)
    kname, pass_bytes, descs = dominant_passes(progs)
    d = descs[0]
    _lib.check(L.tdp_fill(_lib.dptr(T0), 1 << int(d["b_in"]), 7, sh), "fill")
    per = max(1, args.roofline_launches // (4 * len(descs)))
    pass_s = sum(_time_launches(torch, stream, lambda k=k: _lib.check(
        L.tdp_run_pass(_lib.dptr(T0), _lib.dptr(T1), descs[k:k + 1].ctypes.data_as(ctypes.c_void_p), sh),
        "tdp_run_pass"), per) for k in range(len(descs))) / len(descs)
    launches = per * len(descs)
    del T0, T1
    torch.cuda.empty_cache()
    return {"workload": "the %d widest passes of the same banded instance at W=%d, observable 0: 2^%d -> 2^%d "
                        "entries, D=%d, averaged" % (len(descs), W, int(d["b_in"]), int(d["b_out"]), int(d["D"])),
            "kernel": kname, "alg_bytes_per_launch": pass_bytes, "launch_seconds": pass_s,
            "achieved": pass_bytes / pass_s / 1e9, "frac": pass_bytes / pass_s / HBM_PEAK, "launches": launches}


def bench_batched(args, scan):
    from tropidist_amd import batch
    t0 = time.time()
    orders = [list(range(p.ncols)) for p in scan]
    orders, flat, owner = batch.plan_problems(scan, orders)
    plan_s = time.time() - t0
    # one untimed pass over the whole scan (device allocations, code objects), then args.batch_repeats timed
    # passes; the median is reported
    batch.certify_planned(scan, orders, flat, owner)
    import torch
    torch.cuda.synchronize()
    times, ok = [], True
    for _ in range(max(1, args.batch_repeats)):
        t1 = time.perf_counter()
        recs = batch.certify_planned(scan, orders, flat, owner)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
        ok = ok and all(r["witness_ok"] for r in recs if r["distance"] is not None)
    batched_s = sorted(times)[len(times) // 2]
    # plain GPU: the first sample problems one at a time by the single-plan executors (fused passes, fused witness)
    ns = min(args.plain_sample, len(scan))
    sub = [i for i, (pi, _) in enumerate(owner) if pi < ns]
    t2 = time.perf_counter()
    plain = batch.certify_planned_one_at_a_time(scan[:ns], orders[:ns], [flat[i] for i in sub],
                                                [owner[i] for i in sub])
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t2
    same = all(a["per_observable"] == b["per_observable"] and b["witness_ok"] for a, b in zip(recs[:ns], plain))
    widths = sorted(set(pl.peak_bits for pl in flat if pl.feasible))
    graph = scan_graph(flat, torch)
    return {"workload": "%d banded scans, 200-1200 detectors, 5-15 faults/detector, width 10-22, 2 logical rows, "
                        "half with costs 1-3" % len(scan),
            "certs_per_s": len(scan) / batched_s, "batched_seconds": batched_s,
            "plain_gpu_certs_per_s": ns / plain_s, "plain_sample": ns, "plain_seconds": plain_s,
            "speedup_vs_plain": (len(scan) / batched_s) / (ns / plain_s), "witnesses_ok": ok,
            "sample_identical_to_plain": same, "states": int(sum(pl.work() for pl in flat if pl.feasible)),
            "states_per_s": sum(pl.work() for pl in flat if pl.feasible) / batched_s,
            "host_planning_seconds": round(plan_s, 3), "peak_bits_range": [widths[0], widths[-1]],
            "latency_s": batched_s, "repeats_seconds": [round(x, 4) for x in times], "hipgraph": graph}


def scan_graph(flat, torch, repeats=5):
    """The value programs of the scan's wide plans (2^WIDE_BITS entries and more, the multi-program rounds of the
    batched path) run by fused.run_many, against the same schedule captured once in a HIP graph and replayed
    (fused.ProgramGraph): the host time the replays save, median of `repeats` each."""
    from tropidist_amd import batch, fused
    wide = [pl for pl in flat if pl.feasible and pl.peak_bits >= batch.WIDE_BITS]
    if not wide:
        return None
    want = fused.run_many(wide)
    eager = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fused.run_many(wide)
        torch.cuda.synchronize()
        eager.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    g = fused.ProgramGraph(wide)
    capture_s = time.perf_counter() - t0
    g.launch()
    same = g.values() == want
    replay = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        g.launch()
        torch.cuda.synchronize()
        replay.append(time.perf_counter() - t0)
        same = same and g.values() == want
    g.close()
    e, r = sorted(eager)[repeats // 2], sorted(replay)[repeats // 2]
    return {"workload": "value programs of the %d wide plans of the scan (multi-program rounds)" % len(wide),
            "graph_nodes": g.nodes, "run_many_seconds": e, "graph_replay_seconds": r, "capture_seconds": capture_s,
            "saved_seconds": e - r, "saved_per_node_us": 1e6 * (e - r) / max(1, g.nodes), "identical": same}


def _balanced(lo, hi):
    if hi - lo == 1:
        return lo
    mid = (lo + hi) // 2
    return (_balanced(lo, mid), _balanced(mid, hi))


def bench_tables(args):
    """Meet-in-the-middle lookup tables (H rows): every column subset of size <= 3 of the QC code l = m = 12 (288
    columns: 3,981,553 subsets in 2^23 slots at load 0.5) built on the device, best of three builds; then 2^22
    batched lookups of random 3-subsets with a flipped label, timed by HIP events on the kernel's stream.  Beside
    it, the reference's gmitm.c tab_build (OpenMP) on the host for the CPU figure."""
    import ctypes
    import os
    import numpy as np
    import oracle
    from tropidist_amd import _lib, tables, workloads
    t = _lib.torch()
    L = _lib.lib()
    p = workloads.sparse_qc()
    tables.Table(p, 2).free()                   # warm-up
    builds = []
    for _ in range(3):
        tab = tables.Table(p, 3)
        builds.append(tab.build_seconds)
        if len(builds) < 3:
            tab.free()
    entries = int(tab.entries)
    rng = np.random.default_rng(7)
    m = 1 << 22
    # 2^22 independent random 3-subsets (three distinct columns each), not a small set repeated
    idx = rng.integers(0, p.ncols, size=(m, 3))
    while True:
        bad = (idx[:, 0] == idx[:, 1]) | (idx[:, 1] == idx[:, 2]) | (idx[:, 0] == idx[:, 2])
        if not bad.any():
            break
        idx[bad] = rng.integers(0, p.ncols, size=(int(bad.sum()), 3))
    idx.sort(axis=1)
    distinct = len(np.unique(idx[:, 0] * p.ncols * p.ncols + idx[:, 1] * p.ncols + idx[:, 2]))
    zh = tab.zh[idx[:, 0]] ^ tab.zh[idx[:, 1]] ^ tab.zh[idx[:, 2]]
    zl = tab.zl[idx[:, 0]] ^ tab.zl[idx[:, 1]] ^ tab.zl[idx[:, 2]]
    lg = (tab.lg[idx[:, 0]] ^ tab.lg[idx[:, 1]] ^ tab.lg[idx[:, 2]]) ^ np.uint64(1)
    R = np.full(m, 3, dtype=np.int32)
    dev = [t.from_numpy(np.ascontiguousarray(a).view(np.int64 if a.dtype == np.uint64 else np.int32)).cuda()
           for a in (zh, zl, lg, R)]
    res = t.empty(m, dtype=t.int32, device="cuda")
    stream = t.cuda.current_stream()

    def launch():
        _lib.check(L.tdp_tab_lookup_batch(tab.handle, dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(),
                                          dev[3].data_ptr(), m, res.data_ptr(), _lib.stream_handle(stream)),
                   "tdp_tab_lookup_batch")
    look = _time_launches(t, stream, launch, 10)
    hits = int((res >= 0).sum().item())
    out = {"workload": "QC l=m=12 (288 columns), seed 3000: all column subsets of size <= 3 (%d entries, 2^%d slots); "
                       "%d lookups of independent random 3-subsets with a flipped label (%d distinct)"
                       % (entries, tab.log2slots, m, distinct),
           "entries": entries, "build_seconds": min(builds), "ns_per_entry": 1e9 * min(builds) / entries,
           "lookups_per_s": m / look, "lookup_hits": hits, "distinct_queries": distinct,
           "table_bytes": (9 << tab.log2slots),
           "note": "the table's keys and costs (9 bytes per slot) fit in the 256 MB last-level cache (MALL), so "
                   "the probes are served from there, not from HBM"}
    tab.free()
    if not args.no_cpu and oracle.have_ref():
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        c0 = time.perf_counter()
        h, cnt, free = oracle.tab_build(tab.zh, tab.zl, tab.lg, 3, tab.log2slots, 1 if tab.wide else 0, impl="ref")
        cs = time.perf_counter() - c0
        free(h)
        out["cpu_baseline"] = {"value": cnt / cs, "unit": "entries/s", "cores": threads, "kind": "reference",
                               "sample": "the same table built by the reference's gmitm.c tab_build (OpenMP)",
                               "seconds": cs, "entries": int(cnt), "matches_gpu": int(cnt) == entries}
    return out


def bench_search(args):
    """Weight-bounded search (the gmitm route of wide codes): certify the distance 18 of the [[288,12,18]] bivariate
    bicycle code by iterative deepening K = 1..18 -- one anchor per translation orbit, label-free, exclusions, the
    covering bound and a k = 2 table, the reference pipeline's defaults for code instances -- on the device, table
    build included.  Beside it, the reference's gmitm.c gm_search (OpenMP) on the host over the same thresholds,
    whose per-threshold node counts must equal the device's on every exhausted threshold."""
    import os
    import oracle
    from tropidist_amd import search, tables, workloads
    p, rank, anchors = workloads.bivariate_bicycle_288()
    S = search.Searcher(p, rank=rank, anchors=anchors, labelfree=True)
    S.build_exclusions()

    def certify():
        S.tables, S.main = [], -1
        t0 = time.perf_counter()
        tab = tables.Table(p, 2, labelfree=True)
        S.add_table(tab, exact=True)
        runs = []
        for K in range(1, 19):
            r = S.search(K, cover=True)
            runs.append(r)
            if r["best"] is not None or r["status"] != "exhausted":
                break
        dt = time.perf_counter() - t0
        tab.free()
        return dt, runs

    certify()                                  # warm-up: kernels loaded, the search's device buffers allocated
    timed = [certify() for _ in range(3)]
    seconds, runs = sorted(timed, key=lambda x: x[0])[1]
    last = runs[-1]
    nodes = sum(sum(r["nodes"]) for r in runs)
    out = {"workload": "[[288,12,18]] bivariate bicycle code (l=m=12, A=x^3+y^2+y^7, B=y^3+x+x^2), Z basis: distance "
                       "certified by iterative deepening K=1..18, 2 anchors (translation orbits), label-free, "
                       "exclusions, covering bound, k=2 table",
           "distance": last["best"], "status": last["status"], "seconds": seconds, "nodes": nodes,
           "repeats_seconds": [round(t, 4) for t, _ in timed], "timing": "median of 3 certifications after one",
           "nodes_per_s": nodes / seconds, "thresholds": len(runs),
           "witness_ok": bool(last["solutions"]) and search.verify(p, last["solutions"][0]),
           "seconds_per_threshold": [round(r["seconds"], 4) for r in runs],
           "nodes_per_threshold": [sum(r["nodes"]) for r in runs]}
    if not args.no_cpu and oracle.have_gmitm():
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        import numpy as np
        zh, zl = tables.column_hashes(p.h, tables.row_hashes(p.nrows, 12345))
        spec = [(zh, zl, np.zeros(p.ncols, dtype=np.uint64), 2, tables.table_log2slots(p.ncols, 2), 0)]
        c0 = time.perf_counter()
        cpu_runs = []
        for K in range(1, len(runs) + 1):
            cpu_runs.append(oracle.gm_search_ref(S, K, spec, 0, S.rank, S.anchors, use_ex=True, enum_all=False,
                                                 task_depth=max(1, min(4, K - 3)), cover_gate=4))
        cs = time.perf_counter() - c0
        same = all(sum(a["nodes"]) == sum(b["nodes"]) for a, b in zip(cpu_runs[:-1], runs[:-1]))
        out["cpu_baseline"] = {"value": cs, "unit": "seconds (certification)", "cores": threads, "kind": "reference",
                               "sample": "the same thresholds through the reference's gmitm.c gm_search (OpenMP), "
                                         "tables built inside each call",
                               "distance": cpu_runs[-1]["best"], "matches_gpu": cpu_runs[-1]["best"] == last["best"],
                               "exhausted_nodes_identical": same}
        out["speedup_vs_cpu"] = cs / seconds
    return out


def bench_tree(args):
    """Tree-realization DP (T rows): banded R=75 N=300 W=10, one logical row, a balanced tree over the natural
    column order; join pairs per second on the GPU, and the reference's tree_dp.c joins (single-threaded, in the
    same postorder schedule) on the host for the CPU figure."""
    import oracle
    from tropidist_amd import _lib, kernel, problem, trellis, tree_dp
    t = _lib.torch()
    p = problem.banded(R=75, N=300, kmin=2, kmax=5, W=10, nobs=1, rho=0.1, seed=10)
    cols = [h | ((l & 1) << p.nrows) for h, l in zip(p.h, p.l)]
    t0 = time.perf_counter()
    prog = tree_dp.CompiledTreeDP(cols, _balanced(0, len(cols)))
    compile_s = time.perf_counter() - t0
    target = 1 << p.nrows
    lim = dict(max_join_work=1 << 40, max_memory_bytes=1 << 36, max_seconds=600.0)
    value, stats = prog.run(target, **lim)
    t.cuda.synchronize()
    reps = 3
    t1 = time.perf_counter()
    for _ in range(reps):
        value, stats = prog.run(target, **lim)
    t.cuda.synchronize()
    dt = (time.perf_counter() - t1) / reps
    pairs = stats["join_work"]
    dense = kernel.run_plan(trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows))[0]
    out = {"workload": "banded R=75 N=300 k=2-5 W=10, 1 logical row, balanced tree over the natural order, seed 10",
           "distance": value, "dense_distance": dense, "identical": value == dense, "join_pairs": int(pairs),
           "max_join_dimension": stats["max_join_dimension"], "dtype": stats["dtype"],
           "seconds": dt, "pairs_per_s": pairs / dt, "host_compile_seconds": round(compile_s, 3)}
    if not args.no_cpu and oracle.have_ref():
        c0 = time.perf_counter()
        cv = oracle.tree_run(prog, target, impl="ref")
        cs = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": pairs / cs, "unit": "join pairs/s", "cores": 1, "kind": "reference",
                               "sample": "the whole query: the reference's tree_dp.c joins (single-threaded) in the "
                                         "same postorder schedule", "seconds": cs, "distance": cv,
                               "matches_gpu": cv == value}
    out["large"] = bench_tree_large(args)
    return out


def bench_tree_large(args):
    """Tree joins at the scale of the single large instance's merges: banded R=150 N=600 W=13, balanced tree,
    parents of up to 2^26 entries and joins of up to 2^39 pairs.  The widest joins run LDS-tiled (k_tree_join_tiled);
    the bound is the LDS reads, 4 bytes per pair (two dwords per two pairs), against 256 CUs x 128 B/clk x 2.4 GHz.
    The CPU figure is the reference's tree_dp.c on one join of the same tree (the whole query would take about half an
    hour on one core)."""
    import numpy as np
    import oracle
    from tropidist_amd import _lib, kernel, problem, trellis, tree_dp
    t = _lib.torch()
    p = problem.banded(R=150, N=600, kmin=2, kmax=5, W=13, nobs=1, rho=0.1, seed=10)
    cols = [h | ((l & 1) << p.nrows) for h, l in zip(p.h, p.l)]
    prog = tree_dp.CompiledTreeDP(cols, _balanced(0, len(cols)), max_compile_bytes=None)
    target = 1 << p.nrows
    lim = dict(max_join_work=1 << 44, max_memory_bytes=1 << 38, max_seconds=600.0)
    value, stats = prog.run(target, **lim)
    t.cuda.synchronize()
    t1 = time.perf_counter()
    value, stats = prog.run(target, **lim)
    t.cuda.synchronize()
    dt = time.perf_counter() - t1
    pairs = stats["join_work"]
    dense = kernel.run_plan(trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows))[0]
    lds_peak = 256 * 128 * 2.4e9
    out = {"workload": "banded R=150 N=600 k=2-5 W=13, 1 logical row, balanced tree over the natural order, seed 10",
           "distance": value, "dense_distance": dense, "identical": value == dense, "join_pairs": int(pairs),
           "max_join_dimension": stats["max_join_dimension"],
           "max_parent_entries": 1 << max(j.parent_rank for j in prog.joins.values()), "dtype": stats["dtype"],
           "seconds": dt, "pairs_per_s": pairs / dt,
           "roofline": {"bound": "lds", "achieved": 4 * pairs / dt / 1e9, "peak": lds_peak / 1e9, "unit": "GB/s",
                        "frac": 4 * pairs / dt / lds_peak}}
    if not args.no_cpu and oracle.have_ref():
        # one join of about 2^30 pairs, on the host tables of its real size (random finite costs)
        j = min((j for j in prog.joins.values() if j.lmaps), key=lambda j: abs(j.join_dimension - 30))
        rng = np.random.default_rng(1)
        left = rng.integers(0, 20, size=1 << j.left_rank).astype(np.uint16)
        right = rng.integers(0, 20, size=1 << j.right_rank).astype(np.uint16)
        want = np.full(1 << j.parent_rank, prog.infinity, dtype=np.uint16)
        lm, rm = np.array(j.lmaps, dtype=np.uint64), np.array(j.rmaps, dtype=np.uint64)
        c0 = time.perf_counter()
        oracle.tree_join(2, left, right, want, lm, rm, j.parent_rank, j.join_dimension, 0, 1 << j.join_dimension,
                         prog.infinity, impl="ref")
        cs = time.perf_counter() - c0
        L = _lib.lib()
        dl, dr = (t.from_numpy(x.view(np.uint8).copy()).cuda() for x in (left, right))
        do = t.from_numpy(np.full(1 << j.parent_rank, prog.infinity, dtype=np.uint16).view(np.uint8).copy()).cuda()
        _lib.check(L.tdp_tree_join(2, dl.data_ptr(), dr.data_ptr(), do.data_ptr(), lm.ctypes.data, rm.ctypes.data,
                                   j.parent_rank, j.join_dimension, 0, 1 << j.join_dimension, prog.infinity,
                                   _lib.stream_handle()), "tdp_tree_join")
        t.cuda.synchronize()
        g0 = time.perf_counter()
        _lib.check(L.tdp_tree_join(2, dl.data_ptr(), dr.data_ptr(), do.data_ptr(), lm.ctypes.data, rm.ctypes.data,
                                   j.parent_rank, j.join_dimension, 0, 1 << j.join_dimension, prog.infinity,
                                   _lib.stream_handle()), "tdp_tree_join")
        t.cuda.synchronize()
        gs = time.perf_counter() - g0
        same = bool((do.cpu().numpy().view(np.uint16) == want).all())
        out["cpu_baseline"] = {"value": (1 << j.join_dimension) / cs, "unit": "join pairs/s", "cores": 1,
                               "kind": "reference",
                               "sample": "one join of the tree (J=%d, parents 2^%d, children 2^%d and 2^%d) by the "
                                         "reference's tree_join_u16 on random costs 0-19" % (
                                             j.join_dimension, j.parent_rank, j.left_rank, j.right_rank),
                               "seconds": cs, "gpu_seconds_same_join": gs, "matches_gpu": same}
    return out


def _release_cache():
    """Hand torch's cached device blocks back between sections: each section starts from the same free memory
    (the sparse queries alone leave tens of GB cached), and the library's own driver allocations see all of it."""
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def host_cpu(threads=None):
    """The host the CPU figures ran on: logical CPUs (os.cpu_count, as nproc without affinity), the CPUs this process
    may use, the model line of /proc/cpuinfo, and the OpenMP threads of the reference's kernels."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "model": model,
            "omp_threads": threads}


def _omp_threads(n):
    """Set the OpenMP thread count of the reference's kernels (one libgomp in the process); returns the old one."""
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    old = gomp.omp_get_max_threads()
    gomp.omp_set_num_threads(int(n))
    return old


def _ref_values(p):
    """Per-observable values of p by the reference's viterbi.c, driven op by op (the plans of this library's
    planner, whose peak and work equal the reference planner's on the recorded fixtures)."""
    import oracle
    from tropidist_amd import auto, trellis
    order = auto.planning_order(p)
    plans = trellis.build_plans_native(p.h, p.l, [1 << i for i in range(p.nobs)], order, p.nrows, p.w)
    return [oracle.run_plan_c(pl, "ref") if pl.feasible else None for pl in plans]


def bench_cpu_configs(args, scan, requests):
    """The CPU figures of the reference's own kernels beside the GPU ones, for the configurations whose GPU figures
    are reported elsewhere: the plain dense Rank DP on one thread (R=48 N=300 W=10), and certificates per second of
    the batched scan and of the serving mix on all host cores (dense requests by viterbi.c with OpenMP, sparse ones
    by sparse_dp.c under the same sweep logic).  Every CPU sample is bounded to a few seconds per class; values only
    (the reference's traceback is a numpy replay several times slower than its DP, so these rates favour the CPU)."""
    import oracle
    from tropidist_amd import kernel, ordering, problem, sparse_dp, trellis
    if not oracle.have_ref():
        return {"skipped": "oracle/_ref not built"}
    threads = oracle.rlib("viterbi").dp_num_threads()
    out = {"host": host_cpu(threads)}
    # plain dense Rank DP, the reference's CPU path on one thread
    p = problem.banded(R=48, N=300, kmin=1, kmax=4, W=10, nobs=1, rho=0.1, seed=1000)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    gv = kernel.run_plan(plan)[0]
    n, t0 = 0, time.perf_counter()
    while n < 20 or time.perf_counter() - t0 < 0.5:
        kernel.run_plan(plan)
        n += 1
    gpu_s = (time.perf_counter() - t0) / n
    old = _omp_threads(1)
    try:
        m, t0 = 0, time.perf_counter()
        while m < 5 or time.perf_counter() - t0 < 2.0:
            cv = oracle.run_plan_c(plan, "ref")
            m += 1
        cpu_s = (time.perf_counter() - t0) / m
    finally:
        _omp_threads(old)
    out["plain_dense_one_thread"] = {
        "workload": "banded R=48 N=300 k=1-4 W=10, 1 logical row rho=0.1, unit costs, uint8, seed 1000",
        "peak_bits": plan.peak_bits, "work": plan.work(), "distance": gv, "identical": gv == cv,
        "gpu_states_per_s": plan.work() / gpu_s, "gpu_seconds": gpu_s,
        "cpu_baseline": {"value": plan.work() / cpu_s, "unit": "states/s", "cores": 1, "kind": "reference",
                         "sample": "%d runs of the whole plan through viterbi.c on one OpenMP thread" % m,
                         "seconds": cpu_s}}

    def dense_sample(problems, limit_s):
        done, spent = 0, 0.0
        for q in problems:
            t1 = time.perf_counter()
            _ref_values(q)
            spent += time.perf_counter() - t1
            done += 1
            if spent > limit_s:
                break
        return done, spent

    # the batched scan: instances spread evenly through the 512
    if scan:
        pick = [scan[(j * len(scan)) // 24] for j in range(24)]
        done, spent = dense_sample(pick, 8.0)
        out["batched_scan"] = {"cpu_baseline": {
            "value": done / spent, "unit": "certs/s (values of every observable)", "cores": threads,
            "kind": "reference", "sample": "%d of the 512 scan instances, evenly spaced, one at a time through "
                                           "viterbi.c (OpenMP)" % done, "seconds": spent}}
    # the serving mix: each class sampled evenly, weighted by its share
    if requests:
        classes = {}
        for kind, q, K in requests:
            c = "sparse" if kind == "sparse" else ("dense_wide" if getattr(q, "band_width", 0) >= 21 else "dense_narrow")
            classes.setdefault(c, []).append((q, K))
        per, notes = {}, {}
        for c, items in sorted(classes.items()):
            pick = [items[(j * len(items)) // min(12, len(items))] for j in range(min(12, len(items)))]
            if c != "sparse":
                done, spent = dense_sample([q for q, _ in pick], 6.0)
                per[c] = spent / done
                notes[c] = "%d requests" % done
                continue
            done = spent = skipped = 0
            for q, K in pick:
                order = ordering.greedy_order_native(q.h, q.nrows, q.words())
                trs = [sparse_dp.Trellis.for_query(q, i, order) for i in range(q.nobs)]
                if any(t.feasible and t.nslots > 56 for t in trs):
                    skipped += 1                # beyond the reference tables' 56 row slots
                    continue
                t1 = time.perf_counter()
                best = None
                for tr in trs:
                    if not tr.feasible:
                        continue
                    k = K if best is None else best - 1
                    if k < 1:
                        break
                    r = sparse_dp.query(tr, k, mode="mitm", use_h=False, witness=False,
                                        backend=oracle.RefTables(tr))
                    if r["value"] is not None:
                        best = r["value"]
                spent += time.perf_counter() - t1
                done += 1
                if spent > 8.0:
                    break
            per[c] = spent / max(1, done)
            notes[c] = "%d requests through sparse_dp.c (one thread); %d of the sample need more than the " \
                       "reference's 56 row slots and cannot run there" % (done, skipped)
        total = sum(len(v) for v in classes.values())
        rate = 1.0 / sum(len(classes[c]) / total * per[c] for c in per)
        out["serving_mix"] = {"cpu_baseline": {
            "value": rate, "unit": "certs/s (values)", "cores": threads, "kind": "reference",
            "sample": "; ".join("%s: %s, %.1f ms each" % (c, notes[c], 1e3 * per[c]) for c in sorted(per)),
            "class_share": {c: round(len(v) / total, 3) for c, v in sorted(classes.items())}}}
    return out


def bench_sparse(args):
    """cfg3: the QC code l = m = 12 at threshold K (its 77 row slots exceed the reference's 56, so the reference
    cannot run it); beside it, QC l = m = 9 (50 slots) at K = 12 on the GPU and on the reference's sparse_dp.c."""
    import oracle
    from tropidist_amd import ordering, problem, sparse_dp, workloads
    from tropidist_amd.problem import verify_witness
    p = workloads.sparse_qc()
    order = ordering.greedy_straddle_order(p.h, p.nrows)
    tr = sparse_dp.Trellis.for_query(p, 0, order)
    # warm-up at the same threshold: the timed query then finds its tables' memory in torch's caching allocator, as
    # a served query does (a first query in a fresh process also pays the driver's mapping of tens of GB)
    sparse_dp.query(tr, args.sparse_K, mode="mitm", use_h=False, backend="hip", max_states=400_000_000)
    t0 = time.perf_counter()
    r = sparse_dp.query(tr, args.sparse_K, mode="mitm", use_h=False, backend="hip", max_states=400_000_000)
    dt = time.perf_counter() - t0
    steps = sum(r["forward_counts"]) + sum(r["backward_counts"])
    wit = r["witness"]
    out = {"workload": "two-block QC code l=m=12 (288 columns), seed 3000, observable 0, greedy order, "
                       "MITM split K//2+1", "K": args.sparse_K, "value": r["value"], "width": max(tr.width),
           "slots": tr.nslots, "state_steps": int(steps), "state_steps_per_s": steps / dt, "seconds": dt,
           "peak_states": r["peak_states"], "witness_cols": len(wit or []),
           # per state-step: the 16-byte state read, about one 16-byte output written, one random 64-byte line for
           # the partner's probe and one for the output's index insert (the state format here is 16 bytes, the
           # reference's 8)
           "roofline": {"bound": "hbm (random access)", "bytes_per_state_step": 16 + 16 + 64 + 64,
                        "achieved": 160 * steps / dt / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": 160 * steps / dt / HBM_PEAK},
           "witness_ok": bool(wit) and verify_witness(p, sorted(wit), r["value"], 0) if r["value"] else None}
    q = problem.quasi_cyclic(9, 9, seed=3000)
    t2 = sparse_dp.Trellis.for_query(q, 0, ordering.greedy_straddle_order(q.h, q.nrows))
    t0 = time.perf_counter()
    rh = sparse_dp.query(t2, 12, mode="mitm", use_h=False, backend="hip")
    dh = time.perf_counter() - t0
    st2 = sum(rh["forward_counts"]) + sum(rh["backward_counts"])
    cmp_ = {"workload": "QC l=m=9 (162 columns, 50 slots), seed 3000, observable 0, K=12", "value": rh["value"],
            "state_steps": int(st2), "seconds": dh, "state_steps_per_s": st2 / dh}
    if not args.no_cpu and oracle.have_ref():
        t0 = time.perf_counter()
        rr = sparse_dp.query(t2, 12, mode="mitm", use_h=False, backend=oracle.RefTables(t2))
        dr = time.perf_counter() - t0
        cmp_["cpu_baseline"] = {"value": st2 / dr, "unit": "state-steps/s", "cores": 1,
                                "kind": "reference", "sample": "the same query through the reference's sparse_dp.c "
                                "(sp_candidates / sp_merge / sp_join, single-threaded C) under the same sweep logic",
                                "seconds": dr,
                                "matches_gpu": rr["value"] == rh["value"]}
    out["reference_sized"] = cmp_
    return out


def bench_sparse_dense(args):
    """The sparse sweeps' dense fallback (query(dense=f); cdist/sparse_dp.py:221-289, 680-707): a forward sweep at
    K over a banded model of width 24 whose tables fill their 2^width states, plain (sparse tables, the native runner)
    and with dense = 0.25 (dense tables, one tdp_dn_step launch per column); both must give the same value and
    per-cut counts.  The roofline of the dense step comes from HIP events around its launches on the sweep's stream:
    algorithmic bytes per launch = 2^w_old (the old table read once) + 2^w_new (the new one written)."""
    import ctypes

    import numpy as np

    import oracle
    from tropidist_amd import _lib, problem, sparse_dp
    torch = _lib.torch()
    p = problem.banded(R=300, N=2000, kmin=2, kmax=6, W=24, nobs=1, rho=0.05, seed=7024, name="dense_fallback_w24")
    tr = sparse_dp.Trellis.for_query(p, 0, list(range(p.ncols)))
    K = args.dense_fallback_K
    sparse_dp.query(tr, 4, mode="forward", backend="hip", use_h=False, dense=0.25, witness=False)     # warm-up
    runs = {}
    for name, dense in (("plain", None), ("dense", 0.25)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = sparse_dp.query(tr, K, mode="forward", backend="hip", use_h=False, dense=dense, witness=True)
        torch.cuda.synchronize()
        runs[name] = (r, time.perf_counter() - t0)
    plain, dense = runs["plain"][0], runs["dense"][0]
    if plain["value"] != dense["value"] or plain["forward_counts"] != dense["forward_counts"]:
        raise AssertionError("dense fallback differs from the sparse sweep")
    # the same sweep again, recording every dense step's layout; the widest steps are then replayed back to back on
    # fresh tables between two HIP events (each launch alone would also time the host's enqueue gap)
    be = sparse_dp.HipTables(tr)
    launch, rec = be._dense_launch, []

    def recording_launch(*a):
        nst, nend, ncls = a[5], a[8], a[12]
        arrays = [np.frombuffer(ctypes.string_at(a[i], max(1, n) * size), dtype=dt).copy()
                  for i, n, size, dt in ((4, nst, 1, np.uint8), (6, nend, 1, np.uint8), (7, nend, 1, np.uint8),
                                         (13, ncls, 8, np.uint64), (14, ncls, 8, np.int64))]
        rec.append((a, arrays))
        launch(*a)

    be._dense_launch = recording_launch
    sparse_dp.Sweep(tr, +1, K, K, False, backend=be, dense=0.25).run()
    torch.cuda.synchronize()
    top = max(a[3] - a[8] for a, _ in rec)
    wide = [(a, arr) for a, arr in rec if a[3] - a[8] >= top - 1]
    T = torch.randint(0, 12, (1 << max(a[1] for a, _ in wide),), dtype=torch.uint8, device="cuda")
    U = torch.empty(1 << top, dtype=torch.uint8, device="cuda")

    def replay():
        for a, arr in wide:
            args = list(a)
            args[0], args[2] = T.data_ptr(), U.data_ptr()
            args[4], args[6], args[7], args[13], args[14] = (x.ctypes.data for x in arr)
            launch(*args)

    replay()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    replay()
    e1.record()
    torch.cuda.synchronize()
    ksec = e0.elapsed_time(e1) / 1e3
    alg = sum((1 << a[1]) + (1 << (a[3] - a[8])) for a, _ in wide)
    entries = sum(1 << (a[3] - a[8]) for a, _ in rec)
    out = {"workload": "banded R=300 N=2000 k=2-6 W=24 rho=0.05 seed 7024 (25 row slots), observable 0, forward sweep "
                       "at K=%d without lookahead, dense=0.25" % K,
           "value": dense["value"], "dense_cuts": dense["dense_cuts"], "peak_states": dense["peak_states"],
           "plain_seconds": runs["plain"][1], "dense_seconds": runs["dense"][1],
           "speedup_vs_sparse": runs["plain"][1] / runs["dense"][1],
           "dense_entries_per_s": entries / runs["dense"][1],
           "roofline": {"bound": "hbm", "kernel": "tdp::k_dn_step_fast (16 or 4 entries per lane where bits 0-3 or "
                                                   "0-1 pass through)", "achieved": alg / ksec / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": alg / ksec / HBM_PEAK, "launches": len(wide),
                        "widths": [top - 1, top], "alg_bytes_per_launch": alg / len(wide),
                        "launch_seconds": ksec / len(wide),
                        "timing": "the %d widest steps' launches replayed back to back (each call: a counter reset "
                                  "and the step kernel) on fresh tables, HIP events on the stream" % len(wide)},
           "witness_ok": bool(dense["witness"]) and verify_witness_cols(p, dense["witness"], dense["value"])}
    if not args.no_cpu:
        # the oracle's restatement of the reference's numpy Dense steps on the same sweep, stopped after ~12 s
        sw = sparse_dp.Sweep(tr, +1, K, K, False, backend=oracle.OracleTables(tr), dense=0.25)
        cpu = {"entries": 0, "steps": 0, "cut": 0}
        t0 = time.perf_counter()

        def count(c, tab, changed, extra):
            cpu["cut"] = c
            if isinstance(tab, oracle.OracleDense):
                cpu["entries"] += tab.T.size
                cpu["steps"] += 1
            return time.perf_counter() - t0 > 12.0

        sw.run(on_cut=count)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu["entries"] / dt, "unit": "dense entries/s", "cores": 1, "kind": "port",
                               "sample": "the first %d cuts of the same sweep (%d dense steps) through the oracle's "
                                         "numpy restatement of the reference's Dense table" % (cpu["cut"],
                                                                                              cpu["steps"]),
                               "seconds": dt}
    return out


def verify_witness_cols(p, cols, value):
    from tropidist_amd.problem import verify_witness
    return verify_witness(p, sorted(cols), value, 0)


def _serve_run(serve, reqs, rate, seed, max_states):
    # 160 GiB of the 288 GiB HBM: 80 GiB for dense tables and witness arenas, 80 GiB for the resident sparse tables
    # (all sparse requests of the closed-loop run hold about 65 GB at once)
    srv = serve.Server(memory_bytes=160 << 30, dense_lanes=3, sparse_arena_bytes=80 << 30,
                       sparse_max_states=max_states)
    try:
        def note(msg):
            print("bench: serving %s (%.1f s)" % (msg, time.time() - T_START), file=sys.stderr, flush=True)

        if rate is None:                        # closed loop: every request waiting at t = 0
            t0 = time.perf_counter()
            tickets = [srv.submit_dense(*a) if k == "dense" else srv.submit_sparse(*a) for k, a in reqs]
            serve.wait_all(tickets, lambda m: note("closed loop: " + m))
            wall = time.perf_counter() - t0
        else:
            tickets, wall = serve.run_poisson(srv, reqs, rate=rate, seed=seed,
                                              progress=lambda m: note("Poisson stream: " + m))
    finally:
        srv.close()
    return tickets, wall


def bench_serving(args, requests):
    """Capacity first (all requests submitted at once), then a Poisson stream at a fraction of it (or at
    --serving-rate), both on fresh servers and freshly planned requests."""
    from tropidist_amd import _lib, serve
    # the earlier sections leave tens of GB in torch's allocator cache; the server's arenas need that memory
    t = _lib.torch()
    t.cuda.synchronize()
    t.cuda.empty_cache()
    reqs = [("dense", (p,)) if kind == "dense" else ("sparse", (p, K)) for kind, p, K in requests]
    cap_tickets, cap_wall = _serve_run(serve, reqs, None, 0, args.serving_max_states)
    capacity = len(reqs) / cap_wall
    print("bench: serving capacity %.0f/s over %d requests (%.1f s)" % (capacity, len(reqs), time.time() - T_START),
          file=sys.stderr, flush=True)
    rate = args.serving_rate if args.serving_rate else args.serving_load * capacity
    # the Poisson stream: the first --serving-poisson-requests of the same mix (all of them when 0), so that the
    # stream at a fraction of the capacity stays within the run's time
    npois = args.serving_poisson_requests or len(reqs)
    tickets, wall = _serve_run(serve, reqs[:npois], rate, 7, args.serving_max_states)
    print("bench: serving Poisson stream at %.0f/s done (%.1f s)" % (rate, time.time() - T_START), file=sys.stderr,
          flush=True)
    lat = [tk.latency for tk in tickets]
    errs = [tk for tk in tickets if tk.error is not None]
    ok = all(tk.result["witness_ok"] for tk in tickets if tk.error is None and tk.result["distance"] is not None)
    same = all(a.error is None and b.error is None and a.result["distance"] == b.result["distance"]
               for a, b in zip(cap_tickets, tickets) if a.error is None and b.error is None)
    cap_lat = [tk.latency for tk in cap_tickets]
    cap_errs = [tk for tk in cap_tickets if tk.error is not None]
    dl = [tk.latency for tk in tickets if tk.kind == "dense"]
    sl = [tk.latency for tk in tickets if tk.kind == "sparse"]
    # plain GPU: a sample of the same requests answered one at a time, synchronously, by the single-problem paths
    from tropidist_amd import auto
    # the sample is stratified: requests evenly spaced through each class (narrow dense, wide dense, sparse), each
    # class's mean time weighted by its share of the whole mix
    members = {}
    for idx, ((kind, _), (_, p, _)) in enumerate(zip(reqs, requests)):
        c = "sparse" if kind == "sparse" else ("dense_wide" if getattr(p, "band_width", 0) >= 21 else "dense_narrow")
        members.setdefault(c, []).append(idx)
    per_c = max(1, args.serving_plain_sample // max(1, len(members)))
    per_class = {}
    for c, idxs in members.items():
        k = min(per_c, len(idxs))
        for idx in [idxs[(j * len(idxs)) // k] for j in range(k)]:
            kind, a = reqs[idx]
            t1 = time.perf_counter()
            try:
                if kind == "dense":
                    auto.certify(a[0])
                else:
                    serve.sparse_certify(a[0], a[1], max_states=args.serving_max_states)
            except Exception:                   # answered with an error, as the server does
                pass
            per_class.setdefault(c, []).append(time.perf_counter() - t1)
        print("bench: serving plain sample of %s: %d requests, %.1f ms each (%.1f s)"
              % (c, k, 1e3 * sum(per_class[c]) / k, time.time() - T_START), file=sys.stderr, flush=True)
    ns = sum(len(v) for v in per_class.values())
    plain = 1.0 / sum(len(members[c]) / len(reqs) * (sum(v) / len(v)) for c, v in per_class.items())
    return {"workload": "%d requests: 70%% dense width 10-20, 20%% width 21-25, 10%% sparse MITM K 4-%d on QC codes "
                        "l,m 6-15 (tables of at most %d states); capacity with all submitted at once, then the first "
                        "%d of them as Poisson arrivals at %.0f/s (%.0f%% of capacity)" % (
                            len(reqs), args.serving_max_K, args.serving_max_states, len(tickets), rate,
                            100 * rate / capacity),
            "capacity_certs_per_s": capacity, "capacity_wall_seconds": cap_wall, "rate": rate,
            "capacity_p50_s": serve.percentile(cap_lat, 50), "capacity_p99_s": serve.percentile(cap_lat, 99),
            "capacity_errors": len(cap_errs), "poisson_requests": len(tickets),
            "plain_gpu_certs_per_s": plain, "plain_sample": ns, "capacity_vs_plain": capacity / plain,
            "plain_ms_per_request": {c: round(1e3 * sum(v) / len(v), 2) for c, v in sorted(per_class.items())},
            "class_share": {c: round(len(v) / len(reqs), 3) for c, v in sorted(members.items())},
            "certs_per_s": len(tickets) / wall, "wall_seconds": wall, "p50_s": serve.percentile(lat, 50),
            "p99_s": serve.percentile(lat, 99), "dense_p50_s": serve.percentile(dl, 50),
            "dense_p99_s": serve.percentile(dl, 99), "sparse_p50_s": serve.percentile(sl, 50),
            "sparse_p99_s": serve.percentile(sl, 99), "errors": len(errs),
            "certified_per_s": (len(tickets) - len(errs)) / wall, "sparse_max_states": args.serving_max_states,
            "error_kinds": sorted({type(tk.error).__name__ + ": " + str(tk.error)[:80] for tk in errs}),
            "witnesses_ok": ok, "same_distances_both_runs": same}


def profile_step(args):
    """For hardware counters: N headline steps (the four fused programs as multi-program rounds, k_pass_multi), then
    N torch device copies of 2^26 bytes (the reference kernel for scaling the byte counters)."""
    from tropidist_amd import _lib, fused
    torch = _lib.torch()
    L = _lib.lib()
    p, plans, progs, peak, T0, T1, _, _ = dense_setup(L, torch)
    stream = torch.cuda.current_stream()
    for _ in range(args.profile_step):
        values = fused.run_many(plans, stream=stream)
    src, dst = T0[:1 << 26], T1[:1 << 26]
    for _ in range(args.profile_step):
        dst.copy_(src)
    # one more step with per-round HIP events: the rounds' algorithmic bytes (the same launches as the others)
    timing = {}
    fused.run_many(plans, stream=stream, timing=timing)
    torch.cuda.synchronize()
    us, by = timing["round_us"], timing["round_bytes"]
    print(json.dumps({"steps": args.profile_step + 1, "distances": values, "copy_bytes": 1 << 26,
                      "launches_per_step": int(len(by)), "alg_bytes_per_step": float(by.sum()),
                      "alg_bytes_per_launch": float(by.mean()), "event_launch_ns": float(us.mean()) * 1e3}))


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def summarise_step_profile(args):
    """Write profiles/step_pass_multi_banded_w25.json from the three rocprofv3 runs of --profile-step: the kernel
    trace stats, FETCH_SIZE and WRITE_SIZE per dispatch.  The byte counters of k_pass_multi are scaled by what they
    read for the torch copy of 2^26 bytes (the copyBuffer dispatch of the largest grid): scale = bytes moved /
    counter bytes, per counter."""
    trace_csv, fetch_csv, write_csv, log = args.summarise_step_profile
    with open(log) as f:
        info = [json.loads(line) for line in f if line.startswith("{")][-1]
    kernel = "k_pass_round"         # rounds of at most 8 jobs; k_pass_multi runs larger rounds (none in this step)
    stats = [r for r in _csv_rows(trace_csv) if kernel in r["Name"]]
    if not stats:
        raise SystemExit("no %s row in %s" % (kernel, trace_csv))
    # both instantiations (nontemporal input loads on the rounds beyond the memory-side cache, plain otherwise)
    calls = sum(int(r["Calls"]) for r in stats)
    st = {"Name": " + ".join(r["Name"] for r in stats), "Calls": calls,
          "AverageNs": sum(float(r["TotalDurationNs"]) for r in stats) / calls,
          "MinNs": min(int(r["MinNs"]) for r in stats), "MaxNs": max(int(r["MaxNs"]) for r in stats)}

    def counter(path, name):
        rows = [r for r in _csv_rows(path) if r["Counter_Name"] == name]
        mine = [float(r["Counter_Value"]) for r in rows if kernel in r["Kernel_Name"]]
        copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
        g = max(int(r["Grid_Size"]) for r in copies)
        cp = [float(r["Counter_Value"]) for r in copies if int(r["Grid_Size"]) == g]
        return sum(mine) / len(mine), len(mine), sum(cp) / len(cp), len(cp)

    f_k, f_n, f_c, f_cn = counter(fetch_csv, "FETCH_SIZE")
    w_k, w_n, w_c, w_cn = counter(write_csv, "WRITE_SIZE")
    cb = info["copy_bytes"]
    fetch_scale, write_scale = cb / (f_c * 1024), cb / (w_c * 1024)
    rd, wr = f_k * 1024 * fetch_scale, w_k * 1024 * write_scale
    alg = info["alg_bytes_per_launch"]
    out = {"kernel": st["Name"],
           "workload": "the headline step: the four observables' fused programs of the banded R=600 N=6000 W=25 "
                       "instance (seed 2000) as multi-program rounds, one k_pass_round launch per round",
           "alg_bytes_per_launch": alg, "launches_per_step": info["launches_per_step"],
           "alg_bytes_per_step": info["alg_bytes_per_step"],
           "copy_kernel": {"name": "__amd_rocclr_copyBuffer", "bytes_read": cb, "bytes_written": cb,
                           "FETCH_SIZE_KiB": f_c, "WRITE_SIZE_KiB": w_c, "fetch_scale": fetch_scale,
                           "write_scale": write_scale, "dispatches": [f_cn, w_cn]},
           "step_counters": {"FETCH_SIZE_KiB": f_k, "WRITE_SIZE_KiB": w_k, "launches": [f_n, w_n]},
           "traffic_bytes_per_launch": rd + wr, "read_bytes": rd, "write_bytes": wr, "traffic_over_alg": (rd + wr) / alg,
           "trace_calls": int(st["Calls"]), "trace_avg_ns": float(st["AverageNs"]), "trace_min_ns": int(st["MinNs"]),
           "trace_max_ns": int(st["MaxNs"]), "event_avg_ns": info["event_launch_ns"],
           "trace_frac": alg / (float(st["AverageNs"]) * 1e-9) / HBM_PEAK}
    path = os.path.join(ROOT, "profiles", "step_pass_multi_banded_w25.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


def profile_pass(args):
    """Replays for hardware counters: the dominant pass (T0 -> T1) and a torch device copy of 2^b_in bytes."""
    import ctypes
    from tropidist_amd import _lib
    torch = _lib.torch()
    L = _lib.lib()
    p, plans, progs, peak, T0, T1, _, _ = dense_setup(L, torch, W=args.profile_width,
                                                      nobs=None if args.profile_width == 25 else 1)
    kname, alg_bytes, descs = dominant_passes(progs)
    d = descs[0]
    b_in = int(d["b_in"])
    stream = torch.cuda.current_stream()
    sh = _lib.stream_handle(stream)
    _lib.check(L.tdp_fill(_lib.dptr(T0), 1 << b_in, 7, sh), "fill")
    for i in range(args.profile_pass):                  # the dominant passes round-robin
        k = i % len(descs)
        _lib.check(L.tdp_run_pass(_lib.dptr(T0), _lib.dptr(T1), descs[k:k + 1].ctypes.data_as(ctypes.c_void_p), sh),
                   "pass")
    src = T0[:1 << b_in]
    dst = T1[:1 << b_in]
    for _ in range(args.profile_pass):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"pass": {"kernel": kname, "b_in": b_in, "b_out": int(d["b_out"]), "D": int(d["D"]),
                               "threads": 1 << int(d["nfree"]), "alg_bytes": alg_bytes, "passes": len(descs)},
                      "copy_bytes": 1 << b_in, "launches": args.profile_pass}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--roofline-launches", type=int, default=240,
                    help="launches timed for the roofline, spread over the dominant passes")
    ap.add_argument("--gen-workers", type=int, default=16, help="input generator processes (0: in process)")
    ap.add_argument("--skip-dense", action="store_true",
                    help="leave out the headline section (value null; for profiling the other sections alone)")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-sparse", action="store_true")
    ap.add_argument("--skip-sparse-dense", action="store_true", help="skip the sparse sweeps' dense fallback")
    ap.add_argument("--skip-tree", action="store_true")
    ap.add_argument("--skip-tables", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--skip-search", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--batch-repeats", type=int, default=5, help="timed passes over the scan (median reported)")
    ap.add_argument("--plain-sample", type=int, default=512,
                    help="problems of the scan certified one at a time for the plain-GPU rate (default: all)")
    ap.add_argument("--sparse-K", type=int, default=12)
    ap.add_argument("--dense-fallback-K", type=int, default=10)
    ap.add_argument("--serving-requests", type=int, default=10000)
    ap.add_argument("--serving-poisson-requests", type=int, default=2000,
                    help="requests of the Poisson stream, the first ones of the mix (0: all)")
    ap.add_argument("--serving-rate", type=float, default=None,
                    help="Poisson arrival rate (default: --serving-load times the measured capacity)")
    ap.add_argument("--serving-load", type=float, default=0.5,
                    help="Poisson rate as a fraction of the closed-loop capacity")
    ap.add_argument("--serving-plain-sample", type=int, default=60)
    ap.add_argument("--serving-max-K", type=int, default=10)
    ap.add_argument("--serving-max-states", type=int, default=4_000_000,
                    help="state budget of every sparse table, on the served and the one-at-a-time path alike (a "
                         "request past it is answered BudgetExceeded, not certified)")
    ap.add_argument("--profile-pass", type=int, default=0, metavar="N",
                    help="only replay the dominant fused pass N times and a device copy of the same input size N "
                         "times (the reference kernel for scaling hardware byte counters), then exit")
    ap.add_argument("--profile-step", type=int, default=0, metavar="N",
                    help="only run N headline steps and N device copies of 2^26 bytes (for counter runs), then exit")
    ap.add_argument("--summarise-step-profile", nargs=4, metavar=("STATS_CSV", "FETCH_CSV", "WRITE_CSV", "LOG"),
                    help="write the committed counter summary of k_pass_multi from the rocprofv3 outputs of "
                         "--profile-step runs (kernel stats, FETCH_SIZE, WRITE_SIZE) and that run's printed line, "
                         "then exit (no GPU needed)")
    ap.add_argument("--profile-width", type=int, default=25,
                    help="band window of the instance whose widest pass --profile-pass replays (25: the headline, "
                         "2^26 tables; 27: 2^28 tables, larger than the MALL)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the sharding design is written up, "
              "not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.profile_pass:
        profile_pass(args)
        return
    if args.profile_step:
        profile_step(args)
        return
    if args.summarise_step_profile:
        summarise_step_profile(args)
        return
    def progress(msg):
        print("bench: %s (%.1f s)" % (msg, time.time() - T_START), file=sys.stderr, flush=True)

    inputs = generate(args)
    progress("inputs generated")
    from tropidist_amd import _lib
    torch = _lib.torch()
    L = _lib.lib()
    out = {"metric": "Rank-DP states/s (dense trellis, sum of Plan.work / s)", "value": None, "unit": "states/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u8",
           "data": "synthetic (seeded banded detector error models, quasi-cyclic codes)"}
    if not args.skip_dense:
        p, plans, values, dense = bench_dense(args, L, torch)
        out.update(value=dense["value"], ms_per_step=dense["ms_per_step"], config=dense["config"],
                   roofline=dense["roofline"])
        progress("dense done")
    if not args.skip_batched:
        out["batched"] = bench_batched(args, inputs["scan"])
        progress("batched done")
        _release_cache()
    if not args.skip_sparse:
        out["sparse"] = bench_sparse(args)
        progress("sparse done")
        _release_cache()
        if not args.skip_sparse_dense:
            out["sparse"]["dense_fallback"] = bench_sparse_dense(args)
            progress("sparse dense fallback done")
            _release_cache()
    if not args.skip_tree:
        out["tree"] = bench_tree(args)
        progress("tree done")
        _release_cache()
    if not args.skip_tables:
        out["tables"] = bench_tables(args)
        progress("tables done")
    if not args.skip_search:
        out["search"] = bench_search(args)
        progress("search done")
    if not args.skip_serving:
        out["serving"] = bench_serving(args, inputs["requests"])
        progress("serving done")
    if not args.no_cpu:
        out["cpu_configs"] = bench_cpu_configs(args, inputs.get("scan"), inputs.get("requests"))
        progress("CPU baselines done")
    if not args.no_cpu and not args.skip_dense:
        import oracle
        if oracle.have_ref():
            R = oracle.rlib("viterbi")
            threads = R.dp_num_threads()
            c0 = time.perf_counter()
            cv = oracle.run_plan_c(plans[0], "ref")
            cpu_s = time.perf_counter() - c0
            out["cpu_baseline"] = {"value": plans[0].work() / cpu_s, "unit": "states/s", "cores": threads,
                                   "kind": "reference",
                                   "sample": "observable 0 of the same instance (work %d states): the reference's "
                                             "viterbi.c (OpenMP) driven op by op" % plans[0].work(),
                                   "seconds": cpu_s, "distance": cv, "matches_gpu": cv == values[0],
                                   "host": host_cpu(threads)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
