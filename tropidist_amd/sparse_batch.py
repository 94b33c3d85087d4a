"""Many sparse meet-in-the-middle queries resident on one GPU at once (libtdp's tdp_spb_* engine).

    eng = Engine(arena_bytes=4 << 30)
    rid = eng.add(tr, K)                 # one query: a sparse_dp.Trellis, threshold K, default split K // 2 + 1
    while eng.step(): ...                # one column of every active query per call
    status, value, stats = eng.result(rid)
    cols = eng.witness(rid, tr, value); eng.release(rid)

An engine step advances every admitted query by one column: the column steps of all queries are one launch, their
table counts come back in one read, and the new tables' indexes and the forward joins are two more launches.  A query
keeps all its tables on the device until released, so the joins read the backward tables directly and the witness is
traced back on the device in one launch.  Queries can be added between any two steps.

The value of a query is that of sparse_dp.query(tr, K, mode="mitm", use_h=False) (cdist/sparse_dp.py:985-1095 with
the default split): the same column step, caps and join rule; tests/test_gpu_sparse_batch.py checks it against that
path and against the oracle.  certify_requests() answers whole requests (every observable of a problem) with the
records of serve.sparse_certify.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, sparse_dp
from .problem import verify_witness

M64 = (1 << 64) - 1
RUNNING, DONE, STATES, NOMEM, CANCELLED, FAILED = 0, 1, 2, 3, 4, 5


def trellis_words(tr):
    """The engine's view of a trellis: per column (col, fwd mask, fwd bits, bwd mask, bwd bits) as (lo, hi) word
    pairs, shape (N, 10), and the target mask per cut, shape (N + 1, 2)."""
    N = tr.N
    vals = [x for p in range(N) for x in (tr.col[p], tr.fsel_mask[p], tr.fsel_bits[p], tr.bsel_mask[p],
                                          tr.bsel_bits[p])]
    prm = np.zeros((max(N, 1), 10), dtype=np.uint64)
    if N:
        flat = prm.reshape(-1)
        flat[0::2] = [x & M64 for x in vals]
        flat[1::2] = [x >> 64 for x in vals]
    tm = np.zeros((N + 1, 2), dtype=np.uint64)
    tm[:, 0] = [x & M64 for x in tr.tmask]
    tm[:, 1] = [x >> 64 for x in tr.tmask]
    return prm, tm


def query_words(p, order):
    """[(observable, prm, tm)] of the feasible queries of a problem in a column order, built by libtdp's C++ planner
    (tdp_sp_trellis): the words trellis_words gives for sparse_dp.Trellis.for_query(p, obs, order).  ValueError when
    a query needs more row slots than the device states hold, as Trellis raises."""
    from .trellis import parity_targets
    L = _lib.lib()
    N, nq = p.ncols, p.nobs
    if nq == 0:
        return []
    hw = p.words()
    tg = parity_targets(p.l, [1 << i for i in range(nq)])
    prm = np.zeros((nq, max(N, 1), 10), dtype=np.uint64)
    tm = np.zeros((nq, N + 1, 2), dtype=np.uint64)
    ns = np.zeros(nq, dtype=np.int32)
    fe = np.zeros(nq, dtype=np.int32)
    o = np.asarray(order, dtype=np.int32)
    vp = ctypes.c_void_p
    _lib.check(L.tdp_sp_trellis(hw.ctypes.data_as(vp), N, p.nrows, o.ctypes.data_as(vp), tg.ctypes.data_as(vp), nq,
                                sparse_dp.MAX_SLOTS, prm.ctypes.data_as(vp), tm.ctypes.data_as(vp),
                                ns.ctypes.data_as(vp), fe.ctypes.data_as(vp)), "tdp_sp_trellis")
    out = []
    for q in range(nq):
        if not fe[q]:
            continue
        if ns[q] > sparse_dp.MAX_SLOTS:
            raise ValueError("trellis needs %d bits per state (at most %d supported)" % (ns[q], sparse_dp.MAX_SLOTS))
        out.append((q, prm[q], tm[q]))
    return out


class Engine:
    """A tdp_spb handle on one stream.  Not thread-safe on the Python side: one thread drives it."""

    def __init__(self, arena_bytes=4 << 30, max_jobs=4096, stream=None, max_states=30_000_000):
        self.L = _lib.lib()
        self.stream = stream
        self._sh = _lib.stream_handle(stream)
        self.max_states = max_states
        self.max_jobs = max_jobs
        def create():
            h = self.L.tdp_spb_create(arena_bytes, max_jobs, self._sh)
            if not h:
                raise MemoryError(self.L.tdp_last_error().decode(errors="replace"))
            return h

        self.h = _lib.retry_after_release(create)
        self.active = 0
        self._ids = (ctypes.c_int64 * 4096)()

    def close(self):
        if self.h:
            _lib.check(self.L.tdp_spb_free(self.h), "tdp_spb_free")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, tr, K, split=None):
        """Admits the query of a sparse_dp.Trellis; returns its id.  split: the forward share a (default K // 2 + 1,
        as sparse_dp.query)."""
        if not tr.feasible:
            raise ValueError("infeasible trellis: no query to run")
        prm, tm = trellis_words(tr)
        return self.add_words(prm, tm, tr.N, K, split)

    def add_words(self, prm, tm, N, K, split=None):
        """Admits a query given by its words (trellis_words / query_words); returns its id."""
        a = K // 2 + 1 if split is None else int(split)
        a = max(1, min(a, K))
        prm = np.ascontiguousarray(prm, dtype=np.uint64)
        tm = np.ascontiguousarray(tm, dtype=np.uint64)
        assert prm.shape[0] >= N and prm.shape[-1] == 10 and tm.shape == (N + 1, 2)
        rid = ctypes.c_int64(0)
        _lib.check(self.L.tdp_spb_add(self.h, prm.ctypes.data_as(ctypes.c_void_p), tm.ctypes.data_as(ctypes.c_void_p),
                                      N, K, a, self.max_states or 0, ctypes.byref(rid)), "tdp_spb_add")
        self.active += 1
        return rid.value

    def step(self):
        """One column of every active query; returns the number still running."""
        n = ctypes.c_int32(0)
        _lib.check(self.L.tdp_spb_step(self.h, ctypes.byref(n)), "tdp_spb_step")
        self.active = n.value
        return n.value

    def poll(self):
        """Ids of the queries that finished (or failed) since the last poll."""
        out, n = [], ctypes.c_int32(0)
        while True:
            _lib.check(self.L.tdp_spb_poll(self.h, self._ids, len(self._ids), ctypes.byref(n)), "tdp_spb_poll")
            out += self._ids[:n.value]
            if n.value < len(self._ids):
                return out

    def result(self, rid):
        """(status, value or None, {"state_steps", "peak_states", "join_cut", "peak_bytes"})."""
        st, v = ctypes.c_int32(0), ctypes.c_int32(0)
        stats = (ctypes.c_uint64 * 4)()
        _lib.check(self.L.tdp_spb_result(self.h, rid, ctypes.byref(st), ctypes.byref(v), stats), "tdp_spb_result")
        return st.value, (v.value if v.value >= 0 else None), {"state_steps": int(stats[0]),
                                                                "peak_states": int(stats[1]),
                                                                "join_cut": int(stats[2]),
                                                                "peak_bytes": int(stats[3])}

    def stats(self):
        """{"arena_bytes", "in_use", "peak_in_use", "nomem", "steps", "running", "nomem_last"}: "nomem" counts the
        queries that found the arena full, "nomem_last" those of the last step."""
        out = (ctypes.c_uint64 * 7)()
        _lib.check(self.L.tdp_spb_stats(self.h, out), "tdp_spb_stats")
        return dict(zip(("arena_bytes", "in_use", "peak_in_use", "nomem", "steps", "running", "nomem_last"),
                        map(int, out)))

    def witness(self, rid, tr, value):
        """Sorted problem columns of a solution of weight `value` (the query's value); tr: the query's Trellis or
        its column order."""
        order = tr.order if hasattr(tr, "order") else tr
        pos = (ctypes.c_int32 * max(1, value))()
        n = ctypes.c_int32(0)
        _lib.check(self.L.tdp_spb_witness(self.h, rid, pos, max(1, value), ctypes.byref(n)), "tdp_spb_witness")
        return sorted(order[pos[k]] for k in range(n.value))

    def release(self, rid):
        _lib.check(self.L.tdp_spb_release(self.h, rid), "tdp_spb_release")

    def cancel(self, rid):
        """Stops a running query (status CANCELLED, reported by the next poll); no effect on a finished one."""
        _lib.check(self.L.tdp_spb_cancel(self.h, rid), "tdp_spb_cancel")


def threshold_ladder(K, first=4, step=2):
    """Thresholds of the deepening sweeps of a request with threshold K: K, K - step, ... down to `first`, in
    ascending order ([K] when K <= first).  A sweep at k that finds no solution for any observable proves the least
    weight above k, so the first sweep that finds one gives the least weight <= K (cdist/sparse_dp.py:1110-1151
    deepens one unit at a time).  A meet-in-the-middle sweep grows about n-fold per two units of threshold, so the
    lower sweeps cost a small fraction of the last one and end early the requests whose least weight is small."""
    ks = [K]
    while ks[-1] - step >= first:
        ks.append(ks[-1] - step)
    return ks[::-1]


class Request:
    """A whole sparse request: the least weight <= K over every observable of a problem (serve.sparse_certify's
    record).  The request deepens through threshold_ladder(K) (deepen=False: one sweep at K); at each threshold all
    feasible observables run at once, and the first threshold with a solution ends it: the distance is the least
    value and the observable the
    first reaching it -- what the sequential loop with thresholds best - 1 finds.  After a threshold without a
    solution, finish() sets `pending` and the caller admits the request again for the next one."""
    __slots__ = ("problem", "K", "order", "trs", "ids", "record", "error", "tag", "left", "ladder", "stage",
                 "pending", "results")

    def __init__(self, problem, K, order=None, tag=None, deepen=True):
        self.problem, self.K, self.order, self.tag = problem, K, order, tag
        self.trs, self.ids = [], []
        self.record = self.error = None
        self.left = 0
        self.ladder, self.stage, self.pending = (threshold_ladder(K) if deepen else [K]), 0, False
        self.results = {}

    @property
    def k(self):
        """The threshold of the current sweep."""
        return self.ladder[self.stage]

    def plan(self):
        """Column order (the greedy straddle order by default) and the words of each feasible observable's trellis,
        both by the C++ planner (ValueError past the row slot limit, as the single-query path)."""
        from .ordering import greedy_order_native
        p = self.problem
        if self.order is None:
            self.order = greedy_order_native(p.h, p.nrows, p.words())
        self.trs = query_words(p, self.order)
        return self

    def admit(self, eng, owner=None):
        """Adds the request's queries to the engine; owner (a dict id -> request) learns them.  If an add fails, the
        queries added so far stay registered (they drain like the others) and the exception propagates: the caller
        records it as the request's error, which finish() keeps."""
        N = self.problem.ncols
        self.ids = []
        self.results = {}
        self.pending = False
        try:
            for _, prm, tm in self.trs:
                self.ids.append(eng.add_words(prm, tm, N, self.k))
        finally:
            self.left = len(self.ids)
            if owner is not None:
                for rid in self.ids:
                    owner[rid] = self

    def collect(self, eng, rid):
        """Takes the result of a finished query.  A query without a value (or stopped by a budget) needs no witness,
        so its tables are released at once: the arena is free for the others while the request's last queries
        run.  An error here becomes the request's error (finish() keeps it)."""
        try:
            res = eng.result(rid)
            self.results[rid] = res
            if res[1] is None:
                eng.release(rid)
            if res[0] in (NOMEM, STATES, FAILED):  # the request fails: its queries still running stop now
                for other in self.ids:
                    if other not in self.results:
                        eng.cancel(other)
        except Exception as exc:
            if self.error is None:
                self.error = exc

    def finish(self, eng):
        """Builds the record (or the error) once every query has finished, and releases the queries.  An error
        recorded earlier (a failed admission) stands.  A sweep below K without a solution sets `pending` and moves
        to the next threshold instead."""
        try:
            if self.error is not None:
                return self
            best, peak = None, 0
            res = [(i, rid) + tuple(self.results[rid] if rid in self.results else eng.result(rid))
                   for (i, _, _), rid in zip(self.trs, self.ids)]
            for i, rid, st, v, info in res:
                if st == STATES:
                    raise sparse_dp.BudgetExceeded("a table of observable %d exceeded %d states"
                                                   % (i, eng.max_states), "states")
            if any(st == NOMEM for _, _, st, _, _ in res):
                raise MemoryError("sparse batch arena full")
            if any(st == FAILED for _, _, st, _, _ in res):
                raise RuntimeError("a query of the request failed on the device: %s"
                                   % _lib.lib().tdp_last_error().decode(errors="replace"))
            if any(st != DONE for _, _, st, _, _ in res):
                raise RuntimeError("a query of the request was cancelled")
            for i, rid, st, v, info in res:
                peak = max(peak, info.get("peak_states", 0))
                if v is not None and (best is None or v < best[0]):
                    best = (v, i, rid)
            if best is None and self.stage + 1 < len(self.ladder):
                self.stage += 1
                self.pending = True
            elif best is None:
                self.record = {"distance": None, "witness": None, "observable": None, "witness_ok": True,
                               "K": self.K, "path": "engine", "peak_states": peak}
            else:
                v, i, rid = best
                cols = eng.witness(rid, self.order, v)
                self.record = {"distance": v, "witness": cols, "observable": i, "K": self.K,
                               "witness_ok": verify_witness(self.problem, cols, v, i), "path": "engine",
                               "peak_states": peak}
        except Exception as exc:
            self.error = exc
        finally:
            for rid in self.ids:
                if rid not in self.results or self.results[rid][1] is not None:
                    try:
                        eng.release(rid)
                    except Exception as exc:      # reported with the request; the other queries still go
                        if self.error is None:
                            self.error, self.record = exc, None
            self.ids = []
            self.results = {}
        return self


def certify_requests(requests, engine=None, arena_bytes=4 << 30, stream=None, fallback=True):
    """Answers (problem, K[, order]) requests together; returns one Request per input with .record or .error.
    A request whose tables do not fit the arena is answered by the one-at-a-time path (serve.sparse_certify) when
    `fallback` is set."""
    from . import serve
    import collections
    eng = engine or Engine(arena_bytes=arena_bytes, stream=stream)
    try:
        reqs, owner, waiting = [], {}, collections.deque()
        for item in requests:
            r = Request(*item)
            try:
                r.plan()
                if len(r.trs) > eng.max_jobs:       # more queries than the engine has slots
                    if not fallback:
                        raise ValueError("%d queries, the engine holds %d" % (len(r.trs), eng.max_jobs))
                    r.record = serve.sparse_certify(r.problem, r.K, r.order)
                elif r.trs:
                    waiting.append(r)
                else:
                    r.finish(eng)
            except Exception as exc:
                r.error = exc
            reqs.append(r)
        while owner or waiting:
            # admit in arrival order while the engine has job slots for a request's queries
            while waiting and (not owner or len(owner) + len(waiting[0].trs) <= eng.max_jobs):
                r = waiting.popleft()
                try:
                    r.admit(eng, owner)
                except Exception as exc:            # queries already added drain; the request keeps the error
                    r.error = exc
                    if not r.ids:
                        continue
            if not owner:
                continue
            eng.step()
            for rid in eng.poll():
                r = owner.pop(rid, None)
                if r is None:
                    eng.release(rid)
                    continue
                r.collect(eng, rid)
                r.left -= 1
                if r.left:
                    continue
                r.finish(eng)
                if r.pending:                       # no solution at this threshold: the next one
                    waiting.appendleft(r)
                    continue
                if fallback and isinstance(r.error, MemoryError):
                    try:
                        r.error, r.record = None, serve.sparse_certify(r.problem, r.K, r.order)
                    except Exception as exc:
                        r.error = exc
        return reqs
    finally:
        if engine is None:
            eng.close()
