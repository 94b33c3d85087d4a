"""Autotuner of the batched dense executor: workgroups per plan ("group") and checkpoint interval per width bucket.

    tune(widths=range(10, 27), save=True)   times batch.run_plans on synthetic banded plans of each width for the
                                            candidate group sizes and checkpoint intervals, keeps the fastest
    group_for(peak_bits), interval_for(plan)  the tuned choice (or the built-in default when untuned)

The table is cached as JSON (tropidist_amd/autotune_<arch>.json) keyed by width bucket; tuning runs on the GPU in a
few seconds per bucket and is never triggered implicitly."""
from __future__ import annotations

import json
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ARCH = os.environ.get("TDP_ARCH", "gfx950")
CACHE = os.path.join(HERE, "autotune_%s.json" % ARCH)

# built-in defaults, from timings of batch.run_plans on MI355X (banded plans, 16-128 plans per launch)
DEFAULT_GROUPS = {**{b: 1 for b in range(0, 18)}, 18: 4, 19: 4, 20: 4, **{b: 16 for b in range(21, 41)}}
CANDIDATE_GROUPS = (1, 2, 4, 8, 16, 32)
CANDIDATE_INTERVAL_FACTORS = (2, 8, 32)        # S = round(sqrt(f * nops))

_table = None


def _load():
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(CACHE):
            with open(CACHE) as f:
                raw = json.load(f)
            _table = {int(k): v for k, v in raw.get("buckets", {}).items()}
    return _table


def group_for(peak_bits):
    t = _load().get(int(peak_bits))
    if t is not None:
        return int(t["group"])
    return DEFAULT_GROUPS.get(int(peak_bits), 16)


def interval_for(plan):
    t = _load().get(int(plan.peak_bits))
    f = t["interval_factor"] if t is not None else 8
    return max(1, int(round(math.sqrt(f * max(1, plan.nops)))))


def _sample_plans(width, nplans, seed):
    import numpy as np
    from . import batch, workloads
    rng = np.random.default_rng(seed)
    probs = [workloads.banded_instance(rng, seed + i, width, bool(i % 2), nobs=1) for i in range(nplans)]
    _, flat, _ = batch.plan_problems(probs, [list(range(p.ncols)) for p in probs])
    return [p for p in flat if p.feasible]


def tune(widths=range(10, 27), plans_per_width=None, save=True, log=print):
    """Time the candidates per width bucket and (optionally) write the cache.  Returns the table."""
    from . import _lib, batch
    t = _lib.torch()
    out = {}
    for w in widths:
        n = plans_per_width or max(4, min(128, 2 ** max(2, 24 - w)))
        plans = _sample_plans(w, n, 9000 + 100 * w)
        if not plans:
            continue
        b = max(p.peak_bits for p in plans)
        batch.run_plans(plans[:2])                              # warm-up
        best = None
        for g in CANDIDATE_GROUPS:
            t.cuda.synchronize()
            t0 = time.perf_counter()
            batch.run_plans(plans, groups=lambda _: g)
            t.cuda.synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, g)
        g = best[1]
        bestf = None
        for f in CANDIDATE_INTERVAL_FACTORS:
            iv = [max(1, int(round(math.sqrt(f * max(1, p.nops))))) for p in plans]
            t.cuda.synchronize()
            t0 = time.perf_counter()
            batch.run_plans(plans, witness=True, groups=lambda _: g, intervals=iv)
            t.cuda.synchronize()
            dt = time.perf_counter() - t0
            if bestf is None or dt < bestf[0]:
                bestf = (dt, f)
        work = sum(p.work() for p in plans)
        out[b] = {"group": g, "interval_factor": bestf[1], "value_seconds": best[0], "witness_seconds": bestf[0],
                  "plans": len(plans), "states_per_s": work / best[0]}
        if log:
            log("width bucket 2^%d: group %d, interval factor %d, %.3g states/s" % (b, g, bestf[1], work / best[0]))
    if save:
        with open(CACHE, "w") as f:
            json.dump({"arch": ARCH, "buckets": {str(k): v for k, v in out.items()}}, f, indent=1, sort_keys=True)
        global _table
        _table = None
    return out


if __name__ == "__main__":
    tune()
