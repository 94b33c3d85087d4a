"""Certified distance of a pair of matrices on the GPU, with the record of the reference's pipeline.

    certify(p, time_limit=0.0, config=None, log=None)   -> record (the fields of bench/common.py:411-466)
    read_mtx(path) / write_mtx(path, cols, nrows)       MatrixMarket coordinate pattern files (bench/common.py:64-90)

The problem is d = min{|e| : H e = 0, L e != 0}.  The decisions follow the reference's pipeline for a code instance
(bench/adapters/ours_pipeline.py:515-538 with dp=2, the configuration of its reported runs), with this library's
evaluators in place of the C kernels:

  dp >= 1   the dense Rank DP where its table fits the table budget (auto.estimate, ours_pipeline.py:411-426):
            its value is the distance (lower bound), and the fused traceback gives a witness of that weight
            (upper bound), checked against H and L (ours_pipeline.py:458-492).  The reference gives its search a
            head start of the DP's predicted time because its DP runs on one core; here the DP runs first.
  dp == 2   where the dense table does not fit: the threshold sparse DP with meet in the middle, sweeps at the
            certified lower bound l, l + 1, ... (ours_pipeline.py:646-730 without an upper bound): a sweep with no
            solution certifies l = K + 1, one with a solution gives the distance and its witness.  Trellises of
            more than 120 row slots go to the search.
  otherwise the weight-bounded search (search.distance, gm_search on the GPU): thresholds from the lower bound up
            until one yields a solution.

Every bound is recorded with its time in `trace`, as the reference's Trace does (bench/common.py:383-408).  status:
exact (lower bound == upper bound), bounds, timeout, or error (notes says what)."""
from __future__ import annotations

import json
import os
import resource
import time

from . import auto
from .problem import Problem

METHOD = "tropidist_amd"
VERSION = "gfx950"
DEFAULTS = dict(dp=2, kmax=40, kmin=1, k_table=2, table_gb=64)


def configuration(text=""):
    """Settings from "key=value,key=value" (the reference's --config syntax): dp (0, 1 or 2), kmax, kmin, k_table
    and table_gb (device bytes for DP tables, in GiB).  ValueError on unknown keys or values out of range."""
    cfg = dict(DEFAULTS)
    for item in filter(None, (x.strip() for x in text.split(","))):
        if "=" not in item:
            raise ValueError("config item %r is not key=value" % item)
        k, v = (x.strip() for x in item.split("=", 1))
        if k not in DEFAULTS:
            raise ValueError("unknown config key %r (known: %s)" % (k, ", ".join(sorted(DEFAULTS))))
        cfg[k] = int(v)
    if cfg["dp"] not in (0, 1, 2):
        raise ValueError("dp must be 0, 1 or 2")
    for k in ("kmax", "kmin", "k_table"):
        if not 1 <= cfg[k] <= 40:
            raise ValueError(k + " must be between 1 and 40")
    if cfg["kmin"] > cfg["kmax"]:
        raise ValueError("kmin must not exceed kmax")
    if cfg["table_gb"] < 1:
        raise ValueError("table_gb must be at least 1")
    return cfg


# ---------------------------------------------------------------------------------------------- files
def read_mtx(path):
    """(nrows, ncols, cols) of a MatrixMarket "coordinate pattern" file, cols[j] the bitmask of the rows with a one
    in column j (1-based coordinates in the file; a repeated entry cancels, as over GF(2))."""
    with open(path) as f:
        head = f.readline()
        if not head.startswith("%%MatrixMarket matrix coordinate"):
            raise ValueError("%s: not a MatrixMarket coordinate file" % path)
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nrows, ncols, nnz = (int(x) for x in line.split())
        cols = [0] * ncols
        for _ in range(nnz):
            r, c = (int(x) for x in f.readline().split()[:2])
            if not (1 <= r <= nrows and 1 <= c <= ncols):
                raise ValueError("%s: entry (%d, %d) outside %d x %d" % (path, r, c, nrows, ncols))
            cols[c - 1] ^= 1 << (r - 1)
    return nrows, ncols, cols


def write_mtx(path, cols, nrows):
    entries = [(r + 1, j + 1) for j, c in enumerate(cols) for r in range(c.bit_length()) if (c >> r) & 1]
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n")
        f.write("%d %d %d\n" % (nrows, len(cols), len(entries)))
        f.writelines("%d %d\n" % e for e in entries)


def load_matrices(h_path, l_path, name=None):
    """The problem of two MatrixMarket files over the same columns; zero columns and repeated (H, L) columns are
    dropped (they change no distance), with `columns` mapping the kept columns back to the files' indices."""
    nrows, ncols, h = read_mtx(h_path)
    nobs, ncols_l, l = read_mtx(l_path)
    if ncols != ncols_l:
        raise ValueError("H has %d columns and L has %d" % (ncols, ncols_l))
    seen, keep = set(), []
    for j, (a, b) in enumerate(zip(h, l)):
        if (a or b) and (a, b) not in seen:
            seen.add((a, b))
            keep.append(j)
    p = Problem(name or os.path.splitext(os.path.basename(h_path))[0], nrows, nobs, [h[j] for j in keep],
                [l[j] for j in keep])
    p.columns = keep
    p.file_columns = ncols
    return p


# ---------------------------------------------------------------------------------------------- bounds
class Trace:
    """Certified lower and upper bounds with the time each was reached."""

    def __init__(self):
        self.t0 = time.monotonic()
        self.lower = self.upper = None
        self.events = []

    def seconds(self):
        return time.monotonic() - self.t0

    def lb(self, value):
        if self.lower is None or value > self.lower:
            self.lower = value
            self.events.append({"t": round(self.seconds(), 4), "lb": value})

    def ub(self, value):
        if self.upper is None or value < self.upper:
            self.upper = value
            self.events.append({"t": round(self.seconds(), 4), "ub": value})


def _check(p, cols):
    s = lg = 0
    for j in cols:
        s ^= p.h[j]
        lg ^= p.l[j]
    return s == 0 and lg != 0 and len(set(cols)) == len(cols)


# ---------------------------------------------------------------------------------------------- the run
def certify(p, time_limit=0.0, config=None, log=None, truth=None):
    """Record of the certified distance of p (a Problem, unit costs).  time_limit 0 means none."""
    cfg = configuration(config) if isinstance(config, str) or config is None else dict(DEFAULTS, **config)
    say = log or (lambda msg: None)
    tr = Trace()
    deadline = tr.t0 + time_limit if time_limit and time_limit > 0 else None
    notes = {"mode": "code instance", "decisions": []}
    witness, status = None, "bounds"
    tr.lb(cfg["kmin"])

    def remaining():
        return None if deadline is None else deadline - time.monotonic()

    def found(cols):
        nonlocal witness
        cols = sorted(cols)
        if _check(p, cols) and (tr.upper is None or len(cols) < tr.upper):
            witness = cols
            tr.ub(len(cols))

    try:
        budget = cfg["table_gb"] << 30
        est = None
        if cfg["dp"] >= 1:
            est = auto.estimate(p, seeds=2 if p.ncols <= 2000 else 0, time_limit=10.0, max_table_bytes=budget)
            notes["dp"] = {k: est[k] for k in ("cut_rank", "peak_bits", "work", "queries", "fits")}
            notes["dp"]["estimate_seconds"] = round(est["estimate_seconds"], 3)
        if est is not None and est["fits"]:
            notes["decisions"].append("dense DP (table of 2^%d bytes)" % est["peak_bits"])
            say("dense DP: cut rank %d, table 2^%d, work %.3g" % (est["cut_rank"], est["peak_bits"], est["work"]))
            t = time.monotonic()
            value, per = auto.dp_distance(p, est["order"], max_table_bytes=budget, deadline=deadline)
            notes["dp"].update(value=value, per_observable=per, seconds=round(time.monotonic() - t, 3))
            if value is None:
                notes["no_logical"] = True          # no fault set flips a logical observable
            else:
                tr.lb(value)
                obs = per.index(value)
                cols = auto.dp_witness(p, est["order"], obs)
                if cols is not None:
                    found(cols)
        elif cfg["dp"] == 2:
            _sparse_settle(p, cfg, tr, notes, found, deadline, est, say)
        if tr.upper is None and not notes.get("no_logical") and (tr.lower or 1) <= cfg["kmax"]:
            notes["decisions"].append("search from threshold %d" % tr.lower)
            _search(p, cfg, tr, notes, found, remaining, say)
        if tr.upper is not None and tr.lower is not None and tr.lower >= tr.upper:
            status = "exact"
    except TimeoutError as exc:
        status, notes["stopped"] = "timeout", str(exc)
    except MemoryError as exc:
        status, notes["stopped"] = "bounds", "out of memory: %s" % exc
    return _record(p, cfg, tr, status, witness, notes, time_limit, truth)


def _sparse_settle(p, cfg, tr, notes, found, deadline, est, say):
    """Meet-in-the-middle sweeps at the lower bound, then one more each time it rises (ours_pipeline.py:646-730)."""
    info = notes.setdefault("sparse", {"columns": p.ncols})
    order = est["order"] if est is not None else auto.planning_order(p)
    try:
        trs = auto.sparse_trellises(p, order)
    except ValueError as exc:                  # more row slots than a state holds
        info["decision"] = "search (%s)" % exc
        notes["decisions"].append(info["decision"])
        return
    notes["decisions"].append("sparse DP from threshold %d" % tr.lower)
    while tr.upper is None and tr.lower <= cfg["kmax"]:
        K = tr.lower
        t = time.monotonic()
        best = auto.sparse_distance(p, K, order, trellises=trs, lower=K, deadline=deadline,
                                    max_bytes=cfg["table_gb"] << 30, log=None)
        info.setdefault("sweeps", []).append({"K": K, "value": None if best is None else best[0],
                                              "seconds": round(time.monotonic() - t, 3)})
        say("sparse DP K=%d: %s (%.2f s)" % (K, "none" if best is None else best[0], time.monotonic() - t))
        if best is None:
            tr.lb(K + 1)                          # no solution of weight <= K
            continue
        value, cols, _ = best
        tr.lb(value)                              # the sweep returns the least weight <= K
        found(cols)


def _search(p, cfg, tr, notes, found, remaining, say):
    from . import search
    left = remaining()
    if left is not None and left <= 0:
        raise TimeoutError("time limit reached before the search")
    r = search.distance(p, k_table=cfg["k_table"], kmax=cfg["kmax"], start=tr.lower,
                        time_limit=0.0 if left is None else left, log=say)
    notes["search"] = {"k_table": r["k_table"], "table_entries": r["table_entries"],
                       "runs": [{k: run[k] for k in ("K", "status", "best", "seconds")} for run in r["runs"]]}
    tr.lb(r["lower_bound"])
    if r["distance"] is not None and r["solutions"]:
        found(list(r["solutions"][0]))
    if r["runs"] and r["runs"][-1]["status"] == "timeout":
        raise TimeoutError("the search reached the time limit at threshold %d" % r["runs"][-1]["K"])


def _record(p, cfg, tr, status, witness, notes, time_limit, truth):
    cols = getattr(p, "columns", None)
    wit = None if witness is None else ([cols[j] for j in witness] if cols is not None else list(witness))
    rec = {"method": METHOD, "version": VERSION, "config": cfg, "instance": p.name, "variant": "full",
           "rows": p.nrows, "cols": getattr(p, "file_columns", p.ncols), "obs": p.nobs, "threads": 1,
           "time_limit": time_limit, "status": status, "lower_bound": tr.lower, "upper_bound": tr.upper,
           "witness": wit, "witness_ok": None, "seconds": tr.seconds(), "trace": tr.events,
           "peak_rss_mb": round(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024, 1),
           "truth": truth, "truth_kind": "exact" if truth is not None else None, "notes": notes,
           "value_only": False}
    if witness is not None:
        rec["witness_ok"] = _check(p, witness) and len(witness) == tr.upper
        if not rec["witness_ok"]:
            rec["status"] = "error"
            notes["error"] = "witness failed the check against H and L"
    agrees = None
    if truth is not None:
        agrees = not ((tr.lower is not None and tr.lower > truth) or (tr.upper is not None and tr.upper < truth)
                      or (rec["status"] == "exact" and tr.upper not in (None, truth)))
    rec["agrees_with_truth"] = agrees
    return rec


def save(rec, out):
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out + ".tmp", "w") as f:
        json.dump(rec, f, indent=1)
    os.replace(out + ".tmp", out)
