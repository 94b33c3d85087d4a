"""Batched certification of many dense distance problems: narrow plans in persistent GPU launches, wide plans by
fused passes on concurrent streams.

    certify_batch(problems, orders=None)  -> one certificate record per problem (the fields of auto.certify)

Narrow problems (every plan below 2^wide_bits entries): pass 1 runs every (problem, observable) plan for its value;
pass 2 runs, for each problem, the plan of its minimising observable again with the witness traceback.  Each pass
is one launch of tdp_batch_run: a queue of plans sorted by decreasing work, a group of workgroups per plan (no
launch per op, no host round trip per plan).

Wide problems: the multi-program fused executor (fused.run_many) on a stream of its own, while the narrow launches
occupy the device: one launch per round runs the next pass of every wide plan, first the values of every
observable, then the witnesses of the minimising ones.  Witnesses are checked against H and L on the host.

The per-plan checkpoint interval S and the number of resident workgroups come from tropidist_amd.autotune (the
tuned value for the plan's width bucket, or sqrt(8 * ops) and 2 workgroups per CU by default)."""
from __future__ import annotations

import contextlib
import ctypes
import math
import threading
import time

import numpy as np

from . import _lib, auto
from .problem import verify_witness

ALIGN = 256

ITEM_DTYPE = np.dtype([("ops", "<u8"), ("nops", "<i8"), ("peak_bits", "<i4"), ("S", "<i4"),
                       ("tab0", "<u8"), ("tab1", "<u8"), ("ckpt", "<u8"), ("ck_n", "<u8"), ("bits", "<u8"),
                       ("used", "<u8"), ("result", "<u8"), ("barrier", "<u8"), ("witness", "<i4"),
                       ("group", "<i4")])
TICKET_DTYPE = np.dtype([("item", "<i4"), ("member", "<i4")])
# Launches whose plans span several workgroups wait at group barriers, which needs every member resident; one such
# launch runs at a time in the process (launches of single-workgroup plans never wait and run concurrently).
_GROUP_LOCK = threading.Lock()
RESULT_WORDS = 6


def default_group(peak_bits):
    """Workgroups sharing one plan: the autotuner's choice for the width bucket (tropidist_amd.autotune)."""
    from . import autotune
    return autotune.group_for(peak_bits)


def _up(x):
    return (x + ALIGN - 1) // ALIGN * ALIGN


def default_interval(nops):
    return max(1, int(round(math.sqrt(8 * max(1, nops)))))


def plan_bytes(plan, witness, S=None):
    """Device bytes of one batch item (tables, ops, and for witnesses checkpoints, decision maps, used list)."""
    cap = 1 << plan.peak_bits
    nops = plan.nops
    b = 2 * _up(cap) + _up(24 * max(1, nops))
    if witness:
        S = S or default_interval(nops)
        nseg = (nops + S - 1) // S
        b += nseg * _up(cap) + _up(8 * max(1, nseg)) + S * _up(max(4, cap // 8)) + 8 * max(1, nops) + ALIGN
    return b


def run_plans(plans, witness=False, intervals=None, workgroups=None, stream=None, groups=None):
    """Run plans in one persistent launch.  Returns a list of (value or None, used op indices or None).  value 255
    means a cost >= 255 (callers apply final_value).  groups(peak_bits) -> workgroups per plan (default_group)."""
    L = _lib.lib()
    assert ITEM_DTYPE.itemsize == L.tdp_batch_item_bytes(), "BatchItem layout mismatch"
    t = _lib.torch()
    idx = [i for i, p in enumerate(plans) if p.feasible]
    out = [(None, None)] * len(plans)
    if not idx:
        return out
    if intervals is None:
        from . import autotune
        S = {i: autotune.interval_for(plans[i]) for i in idx}
    else:
        S = {i: intervals[i] for i in idx}
    sizes = {i: plan_bytes(plans[i], witness, S[i]) for i in idx}
    arena = _lib.device_empty(sum(sizes.values()) + ALIGN)
    base = _up(arena.data_ptr())
    items = np.zeros(len(idx), dtype=ITEM_DTYPE)
    # queue order: most work first; all op streams first, in one contiguous region copied once
    order = sorted(idx, key=lambda i: -plans[i].work())
    ops_blob = [plans[i].encode().reshape(-1) for i in order]
    ops_off, off = [], 0
    for a in ops_blob:
        ops_off.append(off)
        off += _up(8 * max(3, a.size))
    host = np.zeros(off, dtype=np.uint8)
    for o, a in zip(ops_off, ops_blob):
        host[o:o + a.nbytes] = a.view(np.uint8)
    r0 = base - arena.data_ptr()
    arena[r0:r0 + off].copy_(t.from_numpy(host))
    off += base
    zeroed = t.zeros(len(order) * (8 * RESULT_WORDS + 8), dtype=t.uint8, device="cuda")
    res_base = zeroed.data_ptr()                     # results, then barrier words, of all items (zeroed)
    bar_base = res_base + 8 * RESULT_WORDS * len(order)
    used_base = off                                  # used-op lists of all items, contiguous
    used_off = []
    if witness:
        for i in order:
            used_off.append(off - used_base)
            off += 8 * max(1, plans[i].nops)
        off = _up(off)
    for slot, i in enumerate(order):
        p = plans[i]
        cap = 1 << p.peak_bits
        nops = p.nops
        it = items[slot]
        it["ops"] = base + ops_off[slot]
        it["nops"], it["peak_bits"], it["S"] = nops, p.peak_bits, S[i]
        it["tab0"] = off
        off += _up(cap)
        it["tab1"] = off
        off += _up(cap)
        it["result"] = res_base + 8 * RESULT_WORDS * slot
        it["barrier"] = bar_base + 8 * slot
        it["witness"] = int(witness)
        if witness:
            nseg = (nops + S[i] - 1) // S[i]
            it["ckpt"] = off
            off += nseg * _up(cap)
            it["ck_n"] = off
            off += _up(8 * max(1, nseg))
            it["bits"] = off
            off += S[i] * _up(max(4, cap // 8))
            it["used"] = used_base + used_off[slot]
    assert off <= arena.data_ptr() + arena.numel()
    if workgroups is None:
        cus = ctypes.c_int(0)
        _lib.check(L.tdp_num_cus(ctypes.byref(cus)), "tdp_num_cus")
        workgroups = cus.value                       # one 1024-thread workgroup per CU, all resident
    gfun = groups or default_group
    tickets = []
    for slot, i in enumerate(order):
        g = max(1, min(int(gfun(plans[i].peak_bits)), workgroups))
        items[slot]["group"] = g
        tickets += [(slot, m) for m in range(g)]
    tick = np.array(tickets, dtype=TICKET_DTYPE)
    workgroups = max(1, min(workgroups, len(tick)))
    d_items = t.from_numpy(items.view(np.uint8)).cuda()
    d_tick = t.from_numpy(tick.view(np.uint8)).cuda()
    d_next = t.zeros(1, dtype=t.int32, device="cuda")
    grouped = len(tick) > len(order)
    with (_GROUP_LOCK if grouped else contextlib.nullcontext()):
        _lib.check(L.tdp_batch_run(d_items.data_ptr(), d_tick.data_ptr(), len(tick), d_next.data_ptr(), workgroups,
                                   _lib.stream_handle(stream)), "tdp_batch_run")
        (stream if stream is not None else t.cuda.current_stream()).synchronize()
    res = zeroed[:8 * RESULT_WORDS * len(order)].cpu().numpy().view(np.int64).reshape(-1, RESULT_WORDS)
    used_all = None
    if witness:
        u0 = used_base - arena.data_ptr()
        used_all = arena[u0:u0 + used_off[-1] + 8 * max(1, plans[order[-1]].nops)].cpu().numpy().view(np.int64)
    for slot, i in enumerate(order):
        value, nused, err = (int(x) for x in res[slot][:3])
        if err:
            raise RuntimeError("batched plan %d failed (code %d)" % (i, err))
        used = None
        if witness and value != 255:
            u = used_off[slot] // 8
            used = used_all[u:u + nused].copy()
        out[i] = (value, used)
    return out


def _columns(plan, order, used):
    """Problem columns of the used update ops (op indices -> update ordinal -> column position -> column)."""
    codes = plan.encode()[:, 0]
    ordinal = np.cumsum(codes == 1) - 1
    return sorted(order[plan.columns[int(ordinal[int(i)])]] for i in used)


def plan_problems(problems, orders=None, compile_wide=True):
    """Host planning of a batch with the C++ planner: (orders, flat list of plans, owner (problem, observable) of
    each plan).  compile_wide: also compile the fused programs (value and witness) of plans with 2^WIDE_BITS
    entries or more, which the wide path runs (cached on the plans), so that the device phase does no compiling."""
    from .trellis import build_plans_native
    orders = orders or [auto.planning_order(p) for p in problems]
    flat, owner = [], []
    for pi, (p, o) in enumerate(zip(problems, orders)):
        plans = build_plans_native(p.h, p.l, [1 << i for i in range(p.nobs)], o, p.nrows, p.w, words=p.words())
        for i, plan in enumerate(plans):
            flat.append(plan)
            owner.append((pi, i))
    if compile_wide:
        from . import fused
        for plan in flat:
            if plan.feasible and plan.peak_bits >= WIDE_BITS:
                try:
                    fused.program_for(plan, min_bits=2)
                    fused.witness_program_for(plan)
                except ValueError:          # runs on the per-op executor
                    pass
    return orders, flat, owner


# Problems with a plan of at least 2^WIDE_BITS entries go to the multi-program fused executor; each of its calls
# holds at most WIDE_BUDGET device bytes of tables and witness arenas.
WIDE_BITS = 20
WIDE_BUDGET = 48 << 30


def _chunks(items, cost, budget):
    out, cur, used = [], [], 0
    for it in items:
        c = cost(it)
        if cur and used + c > budget:
            out.append(cur)
            cur, used = [], 0
        cur.append(it)
        used += c
    if cur:
        out.append(cur)
    return out


def _certify_wide(problems, orders, by_problem, pis, budget, stream=None):
    """Records {problem index: record} of the wide problems: the values of every observable's plan, then the
    witnesses of the minimising ones, each as rounds of the multi-program fused executor (fused.run_many); plans
    whose columns exceed the pass limits run alone on the per-op executor."""
    from . import fused, kernel
    recs = {pi: {"distance": None, "per_observable": [None] * problems[pi].nobs, "lower_bound": None,
                 "upper_bound": None, "witness": None, "observable": None, "witness_ok": False} for pi in pis}
    items = [(pi, i, pl) for pi in pis for i, pl in by_problem[pi] if pl.feasible]
    multi = []
    for it in items:
        try:
            fused.program_for(it[2])
            multi.append(it)
        except ValueError:
            recs[it[0]]["per_observable"][it[1]] = kernel.run_plan(it[2], stream=stream, executor="per-op")[0]
    multi.sort(key=lambda it: -it[2].work())
    for chunk in _chunks(multi, lambda it: 2 << it[2].peak_bits, budget):
        for (pi, i, _), v in zip(chunk, fused.run_many([it[2] for it in chunk], stream=stream)):
            recs[pi]["per_observable"][i] = v
    wit = []
    for pi in pis:
        rec = recs[pi]
        vals = [v for v in rec["per_observable"] if v is not None]
        if vals:
            best = min(vals)
            obs = rec["per_observable"].index(best)
            rec.update(distance=best, lower_bound=best, observable=obs)
            wit.append((pi, dict(by_problem[pi])[obs]))

    def settle(pi, v, cols):
        rec = recs[pi]
        ok = v == rec["distance"] and verify_witness(problems[pi], cols, v, rec["observable"])
        rec.update(witness=cols, witness_ok=ok, upper_bound=problems[pi].cost(cols))

    inmulti = []
    for pi, plan in wit:
        try:
            nb = fused.witness_bytes(plan)
        except ValueError:
            nb = None
        if nb is None or nb > budget:
            v, cols = auto.witness_plan(plan, orders[pi], stream=stream,
                                        executor="per-op" if nb is None else "fused")
            settle(pi, v, cols)
        else:
            inmulti.append((pi, plan, nb))
    inmulti.sort(key=lambda t: -t[2])
    for chunk in _chunks(inmulti, lambda t: t[2], budget):
        res = fused.run_many([t[1] for t in chunk], witness=[True] * len(chunk), stream=stream)
        for (pi, plan, _), (v, used) in zip(chunk, res):
            settle(pi, v, _columns(plan, orders[pi], used))
    return recs


def certify_planned(problems, orders, flat, owner, workgroups=None, intervals=None, stream=None,
                    wide_bits=WIDE_BITS, wide_concurrent=True, wide_budget=WIDE_BUDGET, wide_stream=None):
    """The device part of certify_batch on plans made by plan_problems.  Problems with a plan of 2^wide_bits
    entries or more run on the multi-program fused executor (on a stream of their own, alongside the persistent
    launches, when wide_concurrent: wide_stream, or a stream from torch's pool); wide_bits = None keeps every problem
    in the persistent launches."""
    by_problem = {}
    for plan, (pi, i) in zip(flat, owner):
        by_problem.setdefault(pi, []).append((i, plan))
    wide = set()
    if wide_bits is not None:
        wide = {pi for pi, items in by_problem.items()
                if any(pl.feasible and pl.peak_bits >= wide_bits for _, pl in items)}
    recs = [None] * len(problems)
    result, errors, th = {}, [], None
    if wide and wide_concurrent:
        t = _lib.torch()

        def worker():
            try:
                st = wide_stream if wide_stream is not None else t.cuda.Stream()
                with t.cuda.stream(st):
                    result.update(_certify_wide(problems, orders, by_problem, sorted(wide), wide_budget, st))
            except BaseException as e:          # re-raised by the caller
                errors.append(e)

        th = threading.Thread(target=worker, daemon=True)
        th.start()
    try:
        keep = [k for k, (pi, _) in enumerate(owner) if pi not in wide]
        narrow = _certify_narrow(problems, orders, [flat[k] for k in keep], [owner[k] for k in keep],
                                 workgroups, intervals, stream)
        for pi, rec in narrow.items():
            recs[pi] = rec
        if wide and not wide_concurrent:
            result.update(_certify_wide(problems, orders, by_problem, sorted(wide), wide_budget, stream))
    finally:
        if th is not None:
            th.join()
    if errors:
        raise errors[0]
    for pi, rec in result.items():
        recs[pi] = rec
    for pi, p in enumerate(problems):
        if recs[pi] is None:
            recs[pi] = {"distance": None, "per_observable": [None] * p.nobs, "lower_bound": None,
                        "upper_bound": None, "witness": None, "observable": None, "witness_ok": False}
    return recs


def _certify_narrow(problems, orders, flat, owner, workgroups, intervals, stream):
    """Records {problem index: record} of the problems owning `flat`, by the two persistent launches."""
    values = run_plans(flat, witness=False, workgroups=workgroups, stream=stream)
    recs = {pi: {"distance": None, "per_observable": [None] * problems[pi].nobs, "lower_bound": None,
                 "upper_bound": None, "witness": None, "observable": None, "witness_ok": False}
            for pi, _ in owner}
    for (pi, i), (v, _) in zip(owner, values):
        if v is not None:
            if v == 255:
                raise OverflowError("the distance is at least 255, beyond the 8-bit entries of the dense DP")
            recs[pi]["per_observable"][i] = v
    where = {o: k for k, o in enumerate(owner)}
    wplans, wowner = [], []
    for pi, rec in sorted(recs.items()):
        vals = [v for v in rec["per_observable"] if v is not None]
        if vals:
            best = min(vals)
            obs = rec["per_observable"].index(best)
            rec.update(distance=best, lower_bound=best, observable=obs)
            wplans.append(flat[where[(pi, obs)]])
            wowner.append(pi)
    traced = run_plans(wplans, witness=True, workgroups=workgroups, stream=stream,
                       intervals=None if intervals is None else [intervals(p) for p in wplans])
    for pi, plan, (v, used) in zip(wowner, wplans, traced):
        rec = recs[pi]
        cols = _columns(plan, orders[pi], used)
        ok = v == rec["distance"] and verify_witness(problems[pi], cols, v, rec["observable"])
        rec.update(witness=cols, witness_ok=ok, upper_bound=problems[pi].cost(cols))
    return recs


def certify_planned_one_at_a_time(problems, orders, flat, owner):
    """The plain-GPU baseline on the same plans: each problem alone and synchronously, by the single-plan executors
    (kernel.run_plan and auto.witness_plan, fused by default)."""
    from . import kernel
    recs = [{"distance": None, "per_observable": [None] * p.nobs, "witness": None, "witness_ok": False}
            for p in problems]
    by_problem = {}
    for plan, (pi, i) in zip(flat, owner):
        by_problem.setdefault(pi, []).append((i, plan))
    for pi, items in sorted(by_problem.items()):
        rec = recs[pi]
        for i, plan in items:
            rec["per_observable"][i] = kernel.run_plan(plan)[0]
        vals = [v for v in rec["per_observable"] if v is not None]
        if vals:
            best = min(vals)
            obs = rec["per_observable"].index(best)
            plan = dict(items)[obs]
            v, cols = auto.witness_plan(plan, orders[pi])
            rec.update(distance=best, observable=obs, witness=cols,
                       witness_ok=v == best and verify_witness(problems[pi], cols, v, obs))
    return recs


def certify_batch(problems, orders=None, workgroups=None, intervals=None, stream=None):
    """Certificates of many problems: {"distance", "lower_bound", "upper_bound", "witness", "observable",
    "witness_ok", "per_observable"} per problem, as auto.certify; values >= 255 raise OverflowError."""
    t0 = time.time()
    orders, flat, owner = plan_problems(problems, orders)
    t1 = time.time()
    recs = certify_planned(problems, orders, flat, owner, workgroups=workgroups, intervals=intervals, stream=stream)
    t2 = time.time()
    for rec in recs:
        rec["planning_seconds"] = t1 - t0
        rec["batch_seconds"] = t2 - t1
    return recs
