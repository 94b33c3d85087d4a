"""Front end of the exact evaluators on the GPU (the call surface of the reference's cdist/auto.py):

    best_order(p, seeds, time_limit)      the column order to plan with: the natural order and greedy orders
    planning_order(p)                     the natural order (what the batch planners use, without a profile)
    estimate(p, threads, seeds, time_limit, max_table_bytes)
                                          dense DP cost without running anything: table bytes, work, seconds
    dp_distance(p, order)                 exact distance, one GPU plan per logical observable
    dp_witness(p, order, obs)             columns of a minimum-cost solution of one observable, by GPU traceback
    certify(p, order)                     distance + witness + the checks a certificate carries
    sparse_trellises / sparse_estimate / sparse_distance
                                          the weight-truncated sparse DP: predicted cost at a threshold K from
                                          probe sweeps, and the least weight <= K with a witness
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib, kernel, trellis
from .problem import verify_witness

# Table entries touched per second by the fused GPU executor on one MI355X at 2^22-2^26 entries (bench.py, banded
# width-25 instance: about 1e13), taken conservatively; used only for predictions.
RATE_GPU = 5.0e12
# The sparse DP on the GPU: state-steps per second of tdp_sq_run's column steps on large tables and the fixed cost of
# one column step (its launches and the read-back of its counts); predictions only.
SPARSE_STATES_PER_SECOND = 5.0e9
SPARSE_SECONDS_PER_STEP = 6.0e-5
SPARSE_MAX_SLOTS = 120           # row slots of a device state (two words: 64 + 56 bits)


def predicted_seconds(work, threads=None):
    """Seconds for `work` table entries on one GPU (threads is the reference's CPU knob and has no effect)."""
    return work / RATE_GPU


def _order_cost(p, order):
    """(peak cut rank, log2 of the summed table sizes) of a column order: the cost best_order compares."""
    prof = trellis.cut_profile_native(p.h, order, p.nrows, p.words())
    return max(prof, default=0), math.log2(sum(2.0 ** x for x in prof)) if prof else 0.0


def best_order(p, seeds=2, time_limit=None):
    """(order, (peak cut rank, log2 volume)) of the cheapest candidate: the natural order, then `seeds` greedy
    orders (tdp_greedy_order from spread start columns) while time_limit allows.  The reversed natural order of the
    reference's candidates (cdist/auto.py:48-64) has the mirrored cut profile, hence the same cost, and loses the tie,
    so it is not evaluated."""
    t0 = time.time()
    best = (_order_cost(p, list(range(p.ncols))), list(range(p.ncols)))
    if p.ncols:
        from .ordering import greedy_order_native
        for s in range(seeds):
            if time_limit is not None and time.time() - t0 > time_limit:
                break
            o = greedy_order_native(p.h, p.nrows, start=(s * 7919) % p.ncols)
            c = _order_cost(p, o)
            if c < best[0]:
                best = (c, o)
    return best[1], best[0]


def planning_order(p):
    """The natural column order: best_order's choice without greedy candidates (seeds=0), without its profile."""
    return list(range(p.ncols))


def _plans(p, order, observables=None):
    """[(observable, plan)], one dense plan per logical observable sharing H's span form (cdist/auto.py:67-76),
    built by the C++ planner; plans whose tables would exceed 63 active rows come from the Python planner (they are
    only ever measured, never run)."""
    obs = list(range(p.nobs)) if observables is None else list(observables)
    try:
        plans = trellis.build_plans_native(p.h, p.l, [1 << i for i in obs], order, p.nrows, p.w, words=p.words())
    except ValueError:
        hrows = trellis.minimal_span_form(trellis.rows_from_columns(p.h, order, p.nrows))
        pw = None if p.w is None else [p.w[j] for j in order]
        plans = [trellis.make_plan(hrows, trellis.reduce_target_row(trellis.target_row(p.l, 1 << i, order), hrows),
                                   len(order), pw) for i in obs]
    return list(zip(obs, plans))


def estimate(p, threads=1, seeds=2, time_limit=None, max_table_bytes=None, order=None):
    """Dense DP cost of the best order found (or of `order`), with nothing allocated (cdist/auto.py:79-102).

    When the order's cut rank alone already needs a table above max_table_bytes, no plan is built: fits False and
    work None.  Otherwise the first feasible observable's plan stands for all of them: table bytes of its peak, one
    bit more when there are several observables (their target rows may widen a table by one row), and its work times
    the number of observables.  Keys: order, cut_rank, queries, peak_bits, table_bytes, work, predicted_seconds,
    fits, estimate_seconds."""
    t0 = time.time()
    if order is None:
        order, (width, _) = best_order(p, seeds=seeds, time_limit=time_limit)
    else:
        width = max(trellis.cut_profile(p.h, order), default=0)
    rec = {"order": order, "cut_rank": int(width), "queries": p.nobs}
    if max_table_bytes is not None and width > math.log2(max(max_table_bytes, 1)):
        rec.update(peak_bits=int(width), table_bytes=1 << int(width), work=None, predicted_seconds=math.inf,
                   fits=False, estimate_seconds=time.time() - t0)
        return rec
    feasible = [plan for _, plan in _plans(p, order) if plan.feasible]
    bits = feasible[0].peak_bits + (p.nobs > 1) if feasible else 0
    work = feasible[0].work() * p.nobs if feasible else 0
    rec.update(peak_bits=bits, table_bytes=1 << bits, work=work, predicted_seconds=predicted_seconds(work, threads),
               fits=max_table_bytes is None or (1 << bits) <= max_table_bytes, estimate_seconds=time.time() - t0)
    return rec


def dp_distance(p, order, max_table_bytes=None, deadline=None):
    """(least cost over the observables, per-observable costs) of min{cost(e) : H e = 0, (L e)_i = 1}; None where an
    observable has no logical (cdist/auto.py:105-123).  The limits are checked before each observable's plan runs:
    MemoryError when its table would exceed max_table_bytes, TimeoutError past deadline (time.monotonic());
    OverflowError at 255 or more."""
    per = []
    for _, plan in _plans(p, order):
        if deadline is not None and time.monotonic() > deadline:
            raise TimeoutError("DP stopped at the time limit")
        if plan.feasible and max_table_bytes is not None and (1 << plan.peak_bits) > max_table_bytes:
            raise MemoryError("DP table of 2^%d bytes exceeds the limit" % plan.peak_bits)
        per.append(kernel.run_plan(plan)[0] if plan.feasible else None)
    found = [v for v in per if v is not None]
    return (min(found) if found else None), per


# Device bytes the fused witness may use by default (the arena keeps every step's input table).
FUSED_WITNESS_BYTES = 32 << 30


def _columns_of(plan, order, op_indices):
    ops = plan.encode()
    ordinal = np.cumsum(ops[:, 0] == 1) - 1
    return sorted(order[plan.columns[int(ordinal[int(i)])]] for i in op_indices)


def witness_plan(plan, order, max_bytes=None, stream=None, executor="fused"):
    """(value, sorted problem columns of one minimum path) of a feasible plan, by a GPU traceback; None if the
    workspace would exceed max_bytes.  OverflowError at 255.

    executor "fused": the witness-mode fused program (every step's input kept in one arena, a one-wave traceback),
    used when its arena fits max_bytes (default FUSED_WITNESS_BYTES), else the per-op path; "per-op": forward op by
    op with checkpoints every sqrt(ops) ops and segment recomputation (tdp_run_plan_witness)."""
    if executor == "fused":
        from . import fused
        budget = FUSED_WITNESS_BYTES if max_bytes is None else max_bytes
        try:
            fits = fused.witness_bytes(plan) <= budget
        except ValueError:              # a column beyond the pass limits: the per-op witness
            fits = False
        if fits:
            v, used = fused.witness(plan, stream=stream)
            return v, _columns_of(plan, order, used)
    L = _lib.lib()
    nbytes = ctypes.c_uint64(0)
    ops = plan.encode()
    _lib.check(L.tdp_witness_workspace_bytes(len(ops), plan.peak_bits, ctypes.byref(nbytes)), "workspace")
    if max_bytes is not None and nbytes.value > max_bytes:
        return None
    work = _lib.device_empty(nbytes.value)
    used = np.zeros(max(1, len(ops)), dtype=np.int64)
    value, nused = ctypes.c_int(-1), ctypes.c_int64(0)
    rc = L.tdp_run_plan_witness(ops.ctypes.data_as(ctypes.c_void_p), len(ops), plan.peak_bits,
                                _lib.dptr(work), nbytes.value, ctypes.byref(value),
                                used.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nused), _lib.stream_handle(stream))
    _lib.check(rc, "tdp_run_plan_witness")
    v = kernel.final_value(value.value)
    return v, _columns_of(plan, order, used[:nused.value])


def dp_witness(p, order, obs, max_bytes=None):
    """Columns of a minimum-cost solution of min{cost(e) : H e = 0, (L e)_obs = 1} (cdist/auto.py:126-182); None if
    no logical exists or the traceback workspace exceeds max_bytes; OverflowError at 255."""
    plan = _plans(p, order, [obs])[0][1]
    if not plan.feasible:
        return None
    r = witness_plan(plan, order, max_bytes=max_bytes)
    return None if r is None else r[1]


def certify(p, order=None, max_table_bytes=None):
    """Exact distance with a checked witness: {"distance", "lower_bound", "upper_bound", "witness", "observable",
    "witness_ok", "per_observable", "seconds"} (the record fields of bench/common.py:411-466).  The lower bound is
    the DP minimum over all observables; the upper bound is the cost of the witness, checked against H and L."""
    t0 = time.time()
    order = planning_order(p) if order is None else order
    best, per = dp_distance(p, order, max_table_bytes=max_table_bytes)
    rec = {"distance": best, "per_observable": per, "lower_bound": best, "upper_bound": None, "witness": None,
           "observable": None, "witness_ok": False}
    if best is not None:
        obs = per.index(best)
        cols = dp_witness(p, order, obs)
        ok = cols is not None and verify_witness(p, cols, best, obs)
        rec.update(witness=cols, observable=obs, witness_ok=ok, upper_bound=p.cost(cols) if cols else None)
    rec["seconds"] = time.time() - t0
    return rec


# ---------------------------------------------------------------------------------------------- sparse DP
def sparse_trellises(p, order):
    """One sparse trellis per observable of p in the order (infeasible where the observable is a combination of the
    checks); ValueError when one needs more row slots than a device state holds."""
    from . import sparse_dp
    return [sparse_dp.Trellis.for_query(p, i, order) for i in range(p.nobs)]


def sparse_estimate(p, K, order, trellises=None, probe=3, max_seconds=None, max_bytes=None):
    """Predicted cost of one meet-in-the-middle sweep at threshold K per observable, from probe sweeps of the first
    feasible observable (the others differ only in the target row) scaled by their number (cdist/auto.py:201-255).
    The probes keep weights <= 2, and <= `probe` in a direction where that costs little next to the predicted
    sweep, unless the weight-2 prediction already exceeds max_seconds or max_bytes tenfold.  fits False (with a
    reason) when a trellis needs more than SPARSE_MAX_SLOTS slots or no observable can be flipped.  Keys: K, fits,
    queries, slots, cut_rank, schedule, schedule_runs, predicted_work, predicted_seconds, predicted_bytes,
    predicted_peak_states, probe_seconds, probe_weights, estimate_seconds."""
    from . import sparse_dp
    t0 = time.time()
    rec = {"K": K, "fits": False}
    try:
        trs = sparse_trellises(p, order) if trellises is None else trellises
    except ValueError as exc:
        rec.update(reason=str(exc), slots=None, estimate_seconds=time.time() - t0)
        return rec
    live = [t for t in trs if t.feasible]
    rec["queries"] = len(live)
    if not live:
        rec.update(reason="no observable can be flipped", estimate_seconds=time.time() - t0)
        return rec
    rec.update(slots=max(t.nslots for t in live), cut_rank=max(max(t.width) for t in live))
    if rec["slots"] > SPARSE_MAX_SLOTS:
        rec.update(reason="the trellis needs %d row slots, a device state holds %d" % (rec["slots"], SPARSE_MAX_SLOTS),
                   estimate_seconds=time.time() - t0)
        return rec
    steps = 2 * live[0].N * len(live)

    def seconds(info):
        return info["predicted_work"] * len(live) / SPARSE_STATES_PER_SECOND + steps * SPARSE_SECONDS_PER_STEP

    sched, info = sparse_dp.choose_schedule(live[0], K, probe=min(2, probe))
    probe_s = info["probe_seconds"]
    cheap = (max_seconds is None or seconds(info) <= 10 * max_seconds) and \
        (max_bytes is None or info["predicted_bytes"] <= 10 * max_bytes)
    if probe > 2 and cheap:
        budget = min(2e9, 0.25 * info["predicted_work"] * len(live))
        sched, info = sparse_dp.choose_schedule(live[0], K, probe=probe, budget=budget)
        probe_s += info["probe_seconds"]
    rec.update(fits=True, schedule=sched, schedule_runs=info["schedule"],
               predicted_work=info["predicted_work"] * len(live), predicted_seconds=seconds(info),
               predicted_bytes=info["predicted_bytes"], predicted_peak_states=info["predicted_peak_states"],
               probe_seconds=probe_s,
               probe_weights=[len(info["forward_profile"]) - 1, len(info["backward_profile"]) - 1],
               estimate_seconds=time.time() - t0)
    return rec


def sparse_distance(p, K, order, trellises=None, schedule=None, lower=0, deadline=None, max_bytes=None, log=None):
    """(value, columns of p, observable) of the least-weight solution of H e = 0, L e != 0 if its weight is <= K,
    else None (cdist/auto.py:258-284).  One meet-in-the-middle query per observable, each after a solution of
    weight v at threshold v - 1; a solution of weight <= lower (a certified lower bound) ends the search.  schedule:
    a split per cut (e.g. sparse_estimate's), else chosen per observable.  TimeoutError past deadline
    (time.monotonic()); MemoryError where a table would outgrow max_bytes."""
    from . import sparse_dp
    trs = sparse_trellises(p, order) if trellises is None else trellises
    best = None
    for i, tr in enumerate(trs):
        if not tr.feasible:
            continue
        k = K if best is None else best[0] - 1
        if k < max(1, lower):
            break
        split = schedule if schedule is not None and len(schedule) == tr.N + 1 else "schedule"
        if isinstance(split, list):
            split = [min(a, k) for a in split]
        states = None if max_bytes is None else max(1, int(max_bytes // (48 * 8)))
        try:
            r = sparse_dp.query(tr, k, mode="mitm", split=split, use_h=False, witness=True, lower=lower,
                                deadline=deadline, max_states=states, log=log)
        except sparse_dp.BudgetExceeded as exc:
            raise (TimeoutError if exc.kind == "deadline" else MemoryError)(str(exc))
        if r["value"] is not None:
            best = (r["value"], r["witness"], i)
            if r["value"] <= lower:
                break
    return best
