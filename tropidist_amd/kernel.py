"""Executor of a trellis Plan on the GPU (the call surface of the reference's cdist/kernel.py).

    INF, final_value(entry), run_plan(plan, log=None, progress_every=0.0) -> (value | None, stats)

run_plan hands the whole plan to libtdp on one HIP stream; control comes back to Python once per plan, not once per
op as in the reference (cdist/kernel.py:112-135).  Two executors give identical values:
    "fused"   (default) tdp_run_program: per-op kernels while the table is narrow, then passes of up to five
              update columns per kernel (tropidist_amd/fused.py compiles them);
    "per-op"  tdp_run_plan: one streaming kernel per op over the uint8 table in HBM.
run_plan_numpy(plan) keeps the reference's name for its array executor (cdist/kernel.py:72-91) for callers that use
it, and runs the same plan on the device with the per-op kernels; this module has no CPU executor."""
from __future__ import annotations

import ctypes
import time

from . import _lib

INF = 255               # uint8 table entries; additions saturate here


def final_value(entry):
    """Minimum cost from the last table entry of a feasible plan (cdist/kernel.py:18-25): entries are min(true, 255),
    so 255 means a cost of at least 255 and raises OverflowError; never None."""
    v = int(entry)
    if v == INF:
        raise OverflowError("the distance is at least %d, beyond the 8-bit entries of the dense DP" % INF)
    return v


EXECUTORS = ("fused", "per-op")


def run_plan(plan, log=None, progress_every=0.0, stream=None, executor="fused"):
    """Run the plan on the GPU.  Returns (min cost or None if infeasible, stats); OverflowError if the cost is
    >= 255 (cdist/kernel.py:94-138).  `plan` is a Plan or any object with the reference's plan attributes.
    executor: "fused" (passes; a plan of at most 2^tdp_plan_small_bits() entries runs whole in one workgroup's LDS
    instead, "small"), "per-op" (one launch per op) or "small".  stats: seconds, peak_bits, cus (= threads),
    executor."""
    if executor not in EXECUTORS + ("small",):
        raise ValueError("executor must be one of %s" % (EXECUTORS + ("small",),))
    from .trellis import as_plan
    plan = as_plan(plan)
    if not plan.feasible:
        return None, {"feasible": False}
    L = _lib.lib()
    small = plan.peak_bits <= L.tdp_plan_small_bits()
    if executor == "small" and not small:
        raise ValueError("a plan of 2^%d entries is not small (at most 2^%d)" % (plan.peak_bits,
                                                                                L.tdp_plan_small_bits()))
    if executor == "fused" and small:
        executor = "small"
    t0 = time.time()
    if executor == "small":
        ops = plan.encode()
        need = ctypes.c_uint64(0)
        _lib.check(L.tdp_plan_small_bytes(len(ops), ctypes.byref(need)), "tdp_plan_small_bytes")
        work = _lib.device_empty(need.value)
        value = ctypes.c_int(-1)
        _lib.check(L.tdp_run_plan_small(ops.ctypes.data_as(ctypes.c_void_p), len(ops), plan.peak_bits,
                                        work.data_ptr(), need.value, ctypes.byref(value),
                                        _lib.stream_handle(stream)), "tdp_run_plan_small")
        return _finish(L, plan, value.value, executor, t0, log)
    size = 1 << plan.peak_bits
    T0 = _lib.device_empty(size)
    T1 = _lib.device_empty(size)
    if executor == "fused":
        from . import fused
        try:
            fused.program_for(plan)
        except ValueError:              # a column beyond the pass limits: the per-op executor
            executor = "per-op"
    if executor == "fused":
        raw = fused.run_program(plan, None, T0, T1, stream=stream, raw=True)
    else:
        ops = plan.encode()
        value = ctypes.c_int(-1)
        rc = L.tdp_run_plan(ops.ctypes.data_as(ctypes.c_void_p), len(ops), plan.peak_bits, _lib.dptr(T0),
                            _lib.dptr(T1), ctypes.byref(value), _lib.stream_handle(stream))
        _lib.check(rc, "tdp_run_plan")
        raw = value.value
    return _finish(L, plan, raw, executor, t0, log)


def _finish(L, plan, raw, executor, t0, log):
    seconds = time.time() - t0
    if log:
        log("  plan of %d ops on 2^%d entries (%s): %.4f s" % (plan.nops, plan.peak_bits, executor, seconds))
    cus = ctypes.c_int(0)
    _lib.check(L.tdp_num_cus(ctypes.byref(cus)), "tdp_num_cus")
    # "threads" is the key of the reference's stats (its OpenMP thread count); here the device's compute units
    return final_value(raw), {"seconds": seconds, "peak_bits": plan.peak_bits, "cus": cus.value,
                              "threads": cus.value, "executor": executor}


def run_plan_numpy(plan):
    """The value alone, as the reference's run_plan_numpy returns it (None if infeasible, OverflowError at 255),
    computed on the GPU by the per-op executor."""
    return run_plan(plan, executor="per-op")[0]
