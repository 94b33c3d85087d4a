"""Weight-bounded exact search on the GPU: the interface of the reference's cdist/search.py (Searcher, Table,
verify, distance) over the device search of libtdp (tdp_gm_search) and the device tables (tables.Table).

    Searcher(p, rank=None, anchors=None, labelfree=None)
        .build_exclusions()        pairs whose sum is a column never meet in a minimum solution
        .set_kernel(sets)          zero-syndrome zero-logical sets: a minimum solution holds at most half of each
        .add_table(table, zh, zl, exact)
        .search(K, enum_all, task_depth, time_limit, sol_cap, exclusions, pattern, anchors, rank, use_kernel, cover)
        .kernel_sets(wmax_s, colmaps, task_depth, sol_cap)
    distance(p, k_table=2, kmax=20, ...)   iterative deepening from a certified lower bound
    verify(p, cols)

The search tree, its prunings and the records it returns are the reference's (cdist/search.py:145-441,
cdist/gmitm.c:188-1023); the tree is walked on the device, so there is no task list to overflow (status
"too_many_tasks" never occurs) and no task ledger (set_ledger accepts only "off")."""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .tables import Table, column_hashes, row_hashes  # noqa: F401  (the reference's search module exports them)

DMAX = 40               # deepest threshold of the device search

_STATUS = {0: "exhausted", 1: "found", 2: "timeout", 3: "too_many_tasks", 6: "slice_done"}


def set_ledger(path=None, part=(0, 1)):
    """The reference's task ledger (cdist/search.py:36-46) for thresholds that take days: the device search keeps
    none.  Only the default (off) is accepted; anything else raises NotImplementedError."""
    r, m = part
    if not (m >= 1 and 0 <= r < m):
        raise ValueError(f"bad part {part}")
    if path is not None or m != 1:
        raise NotImplementedError("the device search keeps no task ledger and runs no slices")


class _Engine:
    """What the reference's lib() handle is asked for by its callers (gm_num_threads)."""

    def gm_num_threads(self):
        return int(_lib.lib().tdp_gm_num_threads())


def lib():
    return _Engine()


def check_threshold(K):
    """The device search holds one buffer per depth and goes no deeper than DMAX."""
    if K > DMAX:
        raise ValueError(f"threshold {K} exceeds the search depth {DMAX} of the kernel; nothing can be "
                         "concluded at this threshold")


def _ptr(a):
    return None if a is None else a.ctypes.data


def _bits(x):
    out = []
    while x:
        low = x & -x
        out.append(low.bit_length() - 1)
        x ^= low
    return out


def _words(values, nwords):
    """Python-int bitmasks -> (len, nwords) uint64, little-endian words."""
    out = np.zeros((len(values), nwords), dtype=np.uint64)
    for j, v in enumerate(values):
        for w in range(nwords):
            out[j, w] = (v >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
    return out


class Searcher:
    """One problem prepared for the device search: column syndromes as words, the row and column incidence in CSR
    form (rows list their columns in increasing index, the branching order of the search), labels, anchors and
    ranks (cdist/search.py:145-227)."""

    def __init__(self, p, rank=None, anchors=None, labelfree=None):
        self.p = p
        self.labelfree = (p.nobs > 64) if labelfree is None else bool(labelfree)
        n, m = p.ncols, p.nrows
        self.W = max(1, (m + 63) // 64)
        assert self.W < 256
        self.col = _words(p.h, self.W)
        rows_of = [_bits(a) for a in p.h]
        self.wmax = max((len(r) for r in rows_of), default=0)
        self.col_start = np.zeros(n + 1, dtype=np.int32)
        self.col_start[1:] = np.cumsum([len(r) for r in rows_of])
        self.col_rows = np.array([r for rs in rows_of for r in rs] or [0], dtype=np.int32)
        # row incidence: (row, column) pairs sorted by row then column
        pr = np.array([r for rs in rows_of for r in rs], dtype=np.int64)
        pc = np.repeat(np.arange(n, dtype=np.int64), [len(r) for r in rows_of])
        order = np.lexsort((pc, pr))
        self.row_cols = pc[order].astype(np.int32) if order.size else np.zeros(1, dtype=np.int32)
        self.row_start = np.zeros(m + 1, dtype=np.int32)
        self.row_start[1:] = np.cumsum(np.bincount(pr, minlength=m)[:m])
        if self.labelfree:
            self.lw = max(1, (p.nobs + 63) // 64)
            self.full_lg = _words(p.l, self.lw)
            self.lg = np.zeros(n, dtype=np.uint64)
        else:
            self.full_lg = None
            self.lg = np.array([int(x) & 0xFFFFFFFFFFFFFFFF for x in p.l], dtype=np.uint64)
        self.rank = np.ascontiguousarray(rank if rank is not None else np.arange(n), dtype=np.int32)
        self.anchors = np.ascontiguousarray(anchors if anchors is not None else np.arange(n), dtype=np.int32)
        self.tables = []            # (Table, zh, zl) per table, in probe order
        self.main = -1
        self.ex_start = self.ex_cols = None
        self.kernel = None

    def _row(self, r):
        return self.row_cols[self.row_start[r]:self.row_start[r + 1]]

    def build_exclusions(self):
        """Columns x, y sharing a row whose sum (syndrome and label) is a column z: no minimum solution holds two of
        x, y, z, so each forbids the other two below it (cdist/search.py:206-231).  Returns the pair count."""
        p = self.p
        where = {}
        for j, key in enumerate(zip(p.h, p.l)):
            where[key] = j                      # the last column of a repeated (h, l), as the reference keys it
        pairs = set()
        for r in range(p.nrows):
            cs = self._row(r).tolist()
            for a in range(len(cs)):
                for b in range(a + 1, len(cs)):
                    pairs.add((cs[a], cs[b]))
        partners = [set() for _ in range(p.ncols)]
        for x, y in pairs:
            z = where.get((p.h[x] ^ p.h[y], p.l[x] ^ p.l[y]))
            if z is None:
                continue
            for u, v, w in ((x, y, z), (y, x, z), (z, x, y)):
                partners[u].add(v)
                partners[u].add(w)
        lists = [sorted(s) for s in partners]
        self.ex_start = np.zeros(p.ncols + 1, dtype=np.int32)
        self.ex_start[1:] = np.cumsum([len(s) for s in lists])
        flat = [g for s in lists for g in s]
        self.ex_cols = np.array(flat or [0], dtype=np.int32)
        return len(flat) // 2

    def set_kernel(self, sets):
        """Kernel dominance (cdist/search.py:233-254): zero-syndrome zero-logical sets of 3 columns or more; once a
        prefix holds floor(|s|/2) columns of s, the rest of s is forbidden below it.  Returns the set count."""
        sets = [tuple(s) for s in sets if len(s) >= 3]
        for s in sets:
            sy, lg = self.p.check(s)
            assert sy == 0 and lg == 0 and len(set(s)) == len(s), "not a zero-logical kernel element"
        if not sets:
            self.kernel = None
            return 0
        ks_start = np.zeros(len(sets) + 1, dtype=np.int32)
        ks_start[1:] = np.cumsum([len(s) for s in sets])
        ks_cols = np.array([j for s in sets for j in s], dtype=np.int32)
        member = [[] for _ in range(self.p.ncols)]
        for i, s in enumerate(sets):
            for j in s:
                member[j].append(i)
        kc_start = np.zeros(self.p.ncols + 1, dtype=np.int32)
        kc_start[1:] = np.cumsum([len(x) for x in member])
        kc_ids = np.array([i for x in member for i in x] or [0], dtype=np.int32)
        self.kernel = (len(sets), ks_start, ks_cols, kc_start, kc_ids)
        return len(sets)

    def add_table(self, table, zh=None, zl=None, exact=False):
        """Attach a device table: exact=True for the problem's own (a hit is re-derived as a solution), else a
        relaxed problem's table with the per-column hashes (zh, zl) of this problem's columns' images there."""
        assert table.labelfree == self.labelfree, "label-free searches need label-free tables (and vice versa)"
        assert self.labelfree or table.wide or self.p.nobs <= 16, "a compact table compares only 16 observable bits"
        if exact:
            zh, zl = table.zh, table.zl
            self.main = len(self.tables)
        self.tables.append((table, np.ascontiguousarray(zh, dtype=np.uint64),
                            np.ascontiguousarray(zl, dtype=np.uint64)))

    def _call(self, K, tables, main, rank, anchors, use_ex, enum_all, task_depth, time_limit, sol_cap,
              kernel_mode=False, use_kernel=True, cover=False, pattern=None):
        """One device search at threshold K; the record of cdist/search.py:256-330."""
        check_threshold(K)
        L = _lib.lib()
        gate = 0 if not cover else (4 if isinstance(cover, (bool, str)) else max(1, int(round(4 * cover))))
        keep = []
        L.tdp_gm_set_cover(gate, 0 if cover == "simple" else 1)
        if self.kernel is not None and use_kernel:
            nks, a, b, c, d = self.kernel
            L.tdp_gm_set_kernel(nks, _ptr(a), _ptr(b), _ptr(c), _ptr(d))
        else:
            L.tdp_gm_set_kernel(0, None, None, None, None)
        L.tdp_gm_set_kernel_mode(1 if kernel_mode else 0)
        L.tdp_gm_set_full_labels(_ptr(self.full_lg), self.lw if self.full_lg is not None else 0)
        if pattern is not None:
            cls = np.ascontiguousarray(pattern[0], dtype=np.int32)
            onpat = np.ascontiguousarray(pattern[1], dtype=np.uint8)
            keep += [cls, onpat]
            L.tdp_gm_set_pattern(_ptr(cls), _ptr(onpat), len(onpat), int(pattern[2]))
        else:
            L.tdp_gm_set_pattern(None, None, 0, 0)
        nt = len(tables)
        handles = (ctypes.c_void_p * max(nt, 1))(*[t[0].handle for t in tables])
        tzh = (ctypes.c_void_p * max(nt, 1))(*[t[1].ctypes.data for t in tables])
        tzl = (ctypes.c_void_p * max(nt, 1))(*[t[2].ctypes.data for t in tables])
        nodes = np.zeros(DMAX + 2, dtype=np.int64)
        stats = np.zeros(8, dtype=np.int64)
        stride = K + 2
        sol = np.zeros((max(sol_cap, 1), stride), dtype=np.int32)
        best, nsol, status = ctypes.c_int(-1), ctypes.c_int(0), ctypes.c_int(-1)
        t0 = time.time()
        try:
            _lib.check(L.tdp_gm_search(
                self.W, self.p.ncols, self.p.nrows, self.wmax, _ptr(self.col), _ptr(self.lg), _ptr(self.row_start),
                _ptr(self.row_cols), _ptr(self.col_start), _ptr(self.col_rows), nt, handles, tzh, tzl, main,
                _ptr(rank), len(anchors), _ptr(anchors), _ptr(self.ex_start) if use_ex else None,
                _ptr(self.ex_cols) if use_ex else None, K, 1 if enum_all else 0, task_depth, float(time_limit),
                _ptr(nodes), _ptr(stats), ctypes.byref(best), sol_cap, stride, _ptr(sol), ctypes.byref(nsol),
                ctypes.byref(status)), "tdp_gm_search")
        finally:
            L.tdp_gm_set_kernel_mode(0)
            L.tdp_gm_set_kernel(0, None, None, None, None)
            L.tdp_gm_set_cover(0, 0)
            L.tdp_gm_set_full_labels(None, 0)
            L.tdp_gm_set_pattern(None, None, 0, 0)
        dt = time.time() - t0
        if status.value == 4:
            raise RuntimeError(f"the search kernel refused threshold {K} or cut a branch at its depth limit {DMAX}; "
                               "the run is not exhaustive")
        st = _STATUS[status.value]
        sols = [tuple(sorted(int(x) for x in sol[i, 1:1 + int(sol[i, 0])])) for i in range(min(nsol.value, sol_cap))]
        return {
            "K": K, "status": st, "best": None if best.value < 0 else best.value,
            "solutions": sols, "nsol_recorded": nsol.value, "nsol_found": int(stats[7]),
            "solutions_complete": st == "exhausted" and int(stats[7]) == nsol.value,
            "nodes": [int(x) for x in nodes[:K + 2]], "lookups": int(stats[0]),
            "hits": int(stats[1]), "false_hits": int(stats[2]), "tasks": int(stats[3]),
            "cover_pruned": int(stats[4]), "ascent_pruned": int(stats[5]), "cover_calls": int(stats[6]),
            "seconds": dt, "threads": int(L.tdp_gm_num_threads()),
        }

    def kernel_sets(self, wmax_s, colmaps=(), task_depth=2, sol_cap=2000000):
        """Zero-syndrome zero-logical sets of size <= wmax_s reached as canonical prefixes (no smaller such set
        before them), closed under the column maps (cdist/search.py:332-346)."""
        r = self._call(wmax_s, [], -1, self.rank, self.anchors, False, True, max(1, min(task_depth, wmax_s - 1)),
                       0.0, sol_cap, kernel_mode=True, use_kernel=False)
        assert r["solutions_complete"], (r["status"], r["nsol_found"], sol_cap)
        out = {s for s in r["solutions"] if len(s) >= 3}
        if colmaps:
            grown, todo = set(), list(out)
            while todo:
                s = todo.pop()
                if s in grown:
                    continue
                grown.add(s)
                todo.extend(tuple(sorted(cm[j] for j in s)) for cm in colmaps)
            out = grown
        return sorted(out, key=lambda s: (len(s), s)), r

    def search(self, K, enum_all=False, task_depth=None, time_limit=0.0, sol_cap=100000, exclusions=True,
               pattern=None, anchors=None, rank=None, use_kernel=True, cover=False):
        """Every solution of weight <= K (enum_all) or the first one found.  pattern = (cls, onpat, slack): class
        per column (-1: none), on-pattern flag per class, off-pattern columns allowed (cdist/search.py:348-378).
        task_depth only names the depth whose nodes are counted as tasks (default as the reference picks it)."""
        check_threshold(K)
        assert self.main >= 0, "an exact table is required"
        kmain = self.tables[self.main][0].k
        if task_depth is None:
            task_depth = max(1, min(4 if cover else 3, K - kmain - 1))
        a = self.anchors if anchors is None else np.ascontiguousarray(anchors, dtype=np.int32)
        rk = self.rank if rank is None else np.ascontiguousarray(rank, dtype=np.int32)
        use_ex = exclusions and self.ex_start is not None
        return _lib.retry_after_release(lambda: self._call(K, self.tables, self.main, rk, a, use_ex, enum_all,
                                                           task_depth, time_limit, sol_cap, use_kernel=use_kernel,
                                                           cover=cover, pattern=pattern))


def verify(p, cols):
    """Independent check that a column set is an undetected logical fault set (cdist/search.py:381-384)."""
    s, lg = p.check(cols)
    return s == 0 and lg != 0 and len(set(cols)) == len(cols)


def distance(p, k_table=2, kmax=20, relaxed=None, rank=None, anchors=None, enum_all=False, time_limit=0.0, log=None,
             start=1, exclusions=True, on_run=None, table_load=0.5, kernel=None, cover=False, labelfree=None):
    """Exact distance by iterative deepening from `start` (a certified lower bound): thresholds start, start + 1, ...
    until one yields a solution (its weight is the distance) or a run stops short (cdist/search.py:387-441).
    relaxed: (table of a relaxed problem, per-column images of p's columns there) pairs used as pruning tables."""
    S = Searcher(p, rank=rank, anchors=anchors, labelfree=labelfree)
    n_excl = S.build_exclusions() if exclusions else 0
    if kernel:
        S.set_kernel(kernel)
    for tab, img in (relaxed or []):
        zh, zl = tab.hashes_for(img)
        S.add_table(tab, zh, zl)
    main = Table(p, k_table, load=table_load, labelfree=S.labelfree)
    if log:
        log(f"  table k={k_table}: {main.entries} entries in 2^{main.log2slots} slots ({main.build_seconds:.1f}s)")
    S.add_table(main, exact=True)
    out = {"k_table": k_table, "table_entries": main.entries, "table_seconds": main.build_seconds,
           "exclusion_pairs": n_excl, "runs": [], "lower_bound": start, "distance": None, "solutions": [],
           "solutions_complete": False}
    keys = ("K", "status", "best", "nodes", "lookups", "hits", "false_hits", "seconds", "tasks", "cover_pruned",
            "nsol_recorded", "nsol_found", "solutions_complete")
    try:
        K = start
        while K <= kmax:
            r = S.search(K, enum_all=enum_all, time_limit=time_limit, exclusions=exclusions, cover=cover)
            bad = [s for s in r["solutions"] if not verify(p, s)]
            assert not bad, "reported solution failed verification"
            rec = {k: r[k] for k in keys}
            out["runs"].append(rec)
            out["solutions_complete"] = r["solutions_complete"]
            if log:
                log(f"  K={K}: {r['status']} best={r['best']} nodes={sum(r['nodes'])} lookups={r['lookups']} "
                    f"false_hits={r['false_hits']} {r['seconds']:.2f}s")
            if on_run:
                on_run(rec)
            if r["status"] in ("timeout", "too_many_tasks", "slice_done"):
                break
            if r["best"] is not None:
                w = r["best"]
                out.update(distance=w, lower_bound=w, solutions=sorted({s for s in r["solutions"] if len(s) == w}))
                break
            out["lower_bound"] = K + 1
            K += 1
    finally:
        main.free()
    return out
