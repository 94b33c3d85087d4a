"""Front end of the dense exact evaluator on the GPU (the call surface of the reference's cdist/auto.py:43-182):

    best_order(p)              the column order to plan with (the natural order and its cut profile)
    planning_order(p)          the same order without its cut profile (what the planners use)
    estimate(p, ...)           plan cost without running anything: table bytes, work, predicted seconds
    dp_distance(p, order)      exact distance, one GPU plan per logical observable
    dp_witness(p, order, obs)  columns of a minimum-cost solution of one observable, by GPU traceback
    certify(p, order)          distance + witness + the checks a certificate carries

The greedy and local-search column orderings of the reference (cdist/ordering.py) are host heuristics outside this
library; a caller with a better order passes it in."""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib, kernel, trellis
from .problem import verify_witness

# Table entries touched per second by the per-op GPU executor on one MI355X at 2^25-2^26 entries (bench.py,
# banded width-25 instance).  Used only for predictions.
RATE_GPU = 2.4e12


def predicted_seconds(work):
    return work / RATE_GPU


def best_order(p, seeds=0, time_limit=None):
    """(order, (peak cut rank, log2 volume)) of the natural column order, the first candidate of
    cdist/auto.py:48-64.  The second candidate there, the reversed order, has the reversed cut profile (its prefix
    and suffix ranks trade places), hence the same peak and volume, and ties go to the first candidate: it never
    wins, so it is not evaluated."""
    c = list(range(p.ncols))
    prof = trellis.cut_profile_native(p.h, c, p.nrows, p.words())
    return c, (max(prof), math.log2(sum(2.0 ** x for x in prof)))


def planning_order(p):
    """best_order(p)[0] without computing the cut profile: the candidates of cdist/auto.py:48-64 are the natural
    order and its reverse, which ties and loses the tie, so the natural order is always chosen."""
    return list(range(p.ncols))


def _plans(p, order, observables=None):
    """One plan per logical observable (or per listed observable); the span form of H is shared
    (cdist/auto.py:67-76).  Built by the C++ planner, which gives trellis.make_plan's plans exactly."""
    obs = list(range(p.nobs)) if observables is None else list(observables)
    try:
        plans = trellis.build_plans_native(p.h, p.l, [1 << i for i in obs], order, p.nrows, p.w, words=p.words())
    except ValueError:                  # more than 63 active rows: plans too wide for any table, made in Python
        hrows = trellis.minimal_span_form(trellis.rows_from_columns(p.h, order, p.nrows))
        pw = None if p.w is None else [p.w[j] for j in order]
        plans = [trellis.make_plan(hrows, trellis.reduce_target_row(trellis.target_row(p.l, 1 << i, order), hrows),
                                   len(order), pw) for i in obs]
    return list(zip(obs, plans))


def estimate(p, order=None, max_table_bytes=None):
    """DP cost of the order without allocating (cdist/auto.py:79-102): peak bits of the first feasible plan (+1 as a
    margin for the other target rows when there are several), work scaled by the number of observables."""
    t0 = time.time()
    if order is None:
        order, cost = best_order(p)
    else:
        prof = trellis.cut_profile(p.h, order)
        cost = (max(prof), 0.0)
    out = {"order": order, "cut_rank": int(cost[0]), "queries": p.nobs}
    if max_table_bytes is not None and cost[0] > math.log2(max_table_bytes):
        out.update(peak_bits=int(cost[0]), table_bytes=1 << int(cost[0]), work=None, predicted_seconds=math.inf,
                   fits=False, estimate_seconds=time.time() - t0)
        return out
    peak, work = 0, 0
    for _, plan in _plans(p, order):
        if plan.feasible:
            peak, work = plan.peak_bits + (1 if p.nobs > 1 else 0), plan.work() * p.nobs
            break
    out.update(peak_bits=peak, table_bytes=1 << peak, work=work, predicted_seconds=predicted_seconds(work),
               fits=max_table_bytes is None or (1 << peak) <= max_table_bytes, estimate_seconds=time.time() - t0)
    return out


def dp_distance(p, order, max_table_bytes=None, deadline=None):
    """Exact distance min over observables i of min{cost(e) : H e = 0, (L e)_i = 1}, and the per-observable list
    (None where no logical exists), as cdist/auto.py:105-123.  MemoryError if a table would exceed
    max_table_bytes, TimeoutError past deadline (time.monotonic), OverflowError at 255 or more."""
    best, per = None, []
    for i, plan in _plans(p, order):
        if deadline is not None and time.monotonic() > deadline:
            raise TimeoutError("DP stopped at the time limit")
        if not plan.feasible:
            per.append(None)
            continue
        if max_table_bytes is not None and (1 << plan.peak_bits) > max_table_bytes:
            raise MemoryError("DP table of 2^%d bytes exceeds the limit" % plan.peak_bits)
        v, _ = kernel.run_plan(plan)
        per.append(v)
        if v is not None and (best is None or v < best):
            best = v
    return best, per


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
