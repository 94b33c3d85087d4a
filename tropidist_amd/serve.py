"""Serving loop for exact distance queries on one GPU: continuous batching with admission by device footprint.

    srv = Server(memory_bytes=..., dense_lanes=2, sparse_arena_bytes=4 << 30)
    ticket = srv.submit_dense(problem, order=None)          # certificate of the dense DP
    ticket = srv.submit_sparse(problem, K, order=None)      # least weight <= K by the sparse MITM DP, or None
    ticket.wait() -> record;  srv.close()

Dense requests are planned on the host by `planners` threads (the C++ planner and pass compiler release the
interpreter lock), a few requests at a time in arrival order, and then wait in a FIFO queue.  Each dense lane is a
thread with its own HIP stream: when free, it takes the waiting planned requests (a long queue is shared with the
other idle lanes, see claim_count; up to max_batch), admits in arrival order those whose plans fit in the free
device budget, runs them as one batched launch (batch.certify_planned: one persistent kernel for the values, one for
the witnesses) and completes them; requests planned meanwhile are admitted by the next free lane, so batches of
different lanes overlap on the device, and no lane waits for planning.

Sparse requests go to one engine thread with its own stream and a resident sparse_batch.Engine, whose arena
(sparse_arena_bytes, held for the server's life) is set apart from the dense budget.  Between any two engine steps
the thread admits every waiting sparse request (every observable a query of the engine) and completes those whose
queries have finished, so sparse requests that arrive while others run join them on the device, one column of all of
them per step.  Requests are admitted while the arena is less than `sparse_admit` full and no query found it full
at the last step (one at least when nothing runs).  A request one of whose queries finds the arena full stops its
other queries and goes back to the head of the queue while others run.  A request that does not fit alone, or has
more queries than the engine has slots, goes to a lane of its own with its own stream, where the one-at-a-time path
(sparse_certify, checkpointed sweeps) answers it while the engine keeps stepping the others; so does a request whose
trellis needs more row slots than a device state holds, which sparse_certify answers by the weight-bounded search.
A sparse request sweeps at its threshold K, as the one-at-a-time path does (sparse_deepen=True deepens through
sparse_batch.threshold_ladder(K) first: on the full serving mix with K <= 10 that took 33.7 s for its 1,000 sparse
requests against 22.9 s without, and let two requests finish below the state budget that the one-at-a-time path
refuses).  Every table is held to `sparse_max_states` states on both paths; a request past that budget is answered
with sparse_dp.BudgetExceeded.

The result of a request does not depend on what it was batched with: every dense plan runs alone in its workgroup on
its own tables, every sparse query on its own tables and counters (tests/test_gpu_serve.py checks each result alone
and in the mix)."""
from __future__ import annotations

import collections
import threading
import time

from . import _lib, auto, batch, sparse_batch, sparse_dp


class Ticket:
    __slots__ = ("kind", "problem", "order", "K", "t_submit", "t_start", "t_done", "result", "error", "_event",
                 "plans", "footprint")

    def __init__(self, kind, problem, order, K):
        self.kind, self.problem, self.order, self.K = kind, problem, order, K
        self.t_submit = time.perf_counter()
        self.t_start = self.t_done = None
        self.result = self.error = None
        self._event = threading.Event()
        self.plans = None
        self.footprint = 0

    def wait(self, timeout=None):
        """The request's record; raises its error, or TimeoutError when it is not answered within `timeout`."""
        if not self._event.wait(timeout):
            raise TimeoutError("request not answered within %s s" % timeout)
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


def _dense_footprint(plans):
    """Device bytes a dense request may hold at once: persistent-launch items for narrow plans; for a wide request
    (a plan of 2^batch.WIDE_BITS entries or more) the fused tables of every plan plus the largest witness arena."""
    from . import fused
    if any(p.peak_bits >= batch.WIDE_BITS for p in plans):
        wit = 0
        for p in plans:
            try:
                wit = max(wit, fused.witness_bytes(p))
            except ValueError:
                wit = max(wit, batch.plan_bytes(p, True))
        return sum(2 << p.peak_bits for p in plans) + wit
    return sum(batch.plan_bytes(p, False) for p in plans) + max((batch.plan_bytes(p, True) for p in plans), default=0)


class Server:
    def __init__(self, memory_bytes=64 << 30, dense_lanes=2, max_batch=1024, sparse_arena_bytes=4 << 30,
                 sparse_max_states=30_000_000, sparse_admit=0.5, sparse_max_jobs=4096, prewarm=True,
                 sparse_deepen=False, planners=4):
        if sparse_arena_bytes >= memory_bytes:
            raise ValueError("the sparse arena (%d bytes) must leave room for dense requests in %d bytes"
                             % (sparse_arena_bytes, memory_bytes))
        self.budget = memory_bytes - sparse_arena_bytes
        self.sparse_admit = sparse_admit        # admit sparse requests while the arena is less full than this
        self.sparse_deepen = sparse_deepen
        self.in_use = 0
        self.max_batch = max_batch
        self.idle_lanes = 0                     # dense lanes waiting for requests
        self.share_above = 64                   # queue length from which a lane leaves a share to idle lanes
        self.plan_q = collections.deque()       # dense requests waiting for host planning
        self.planning = 0                       # dense requests being planned
        self.dense_q = collections.deque()      # planned dense requests waiting for a lane
        self.sparse_q = collections.deque()
        self.sparse_failed = None               # the error that stopped the engine thread, if any
        self.cv = threading.Condition()
        self.closed = False
        self.done = []
        _lib.lib()
        t = _lib.torch()
        # Each dense lane keeps two streams for its life: its persistent launches, and the wide problems' rounds
        # that run beside them.  torch's caching allocator keeps freed blocks per stream, so fixed streams let a
        # lane's tables, workspaces and witness arenas reuse its own cached blocks.  prewarm: one segment per stream
        # (the lane's share of the dense budget for the wide stream, a quarter of that for the other) is taken and
        # given back here, so that those allocations split cached memory instead of calling hipMalloc while requests
        # run (a traced serving run showed single hipMalloc calls of 2 s, during which the other lanes' launches
        # waited).
        self.threads = []
        # host planning (the C++ planner and pass compiler, which release the interpreter lock) runs on threads of
        # its own, so that a lane's device work never waits for the planning of the requests it claims
        self.planner_threads = []
        for _ in range(max(1, planners)):
            th = threading.Thread(target=self._planner, daemon=True)
            th.start()
            self.planner_threads.append(th)
        share = self.budget // max(1, dense_lanes)
        for k in range(dense_lanes):
            st_main, st_wide = t.cuda.Stream(), t.cuda.Stream()
            if prewarm:
                for st, nbytes in ((st_wide, share), (st_main, share // 4)):
                    with t.cuda.stream(st):
                        block = t.empty(max(1, nbytes), dtype=t.uint8, device="cuda")
                        del block
            th = threading.Thread(target=self._dense_lane, args=(st_main, st_wide), daemon=True)
            th.start()
            self.threads.append(th)
        sst = t.cuda.Stream()
        self.sparse_max_states = sparse_max_states
        self.sparse_engine = sparse_batch.Engine(arena_bytes=sparse_arena_bytes, stream=sst,
                                                 max_states=sparse_max_states, max_jobs=sparse_max_jobs)
        self.engine_thread = threading.Thread(target=self._sparse_engine_loop, args=(sst,), daemon=True)
        self.engine_thread.start()
        # requests the engine cannot hold (more queries than its slots, or tables beyond its arena alone) are
        # answered one query at a time on a lane of their own, so that the engine keeps stepping the others
        self.single_q = collections.deque()
        self.single_closed = False
        self.single_thread = threading.Thread(target=self._sparse_single_lane, daemon=True)
        self.single_thread.start()

    # ------------------------------------------------------------------ submission
    def submit_dense(self, problem, order=None):
        tk = Ticket("dense", problem, order, None)
        with self.cv:
            self.plan_q.append(tk)
            self.cv.notify_all()
        return tk

    def submit_sparse(self, problem, K, order=None):
        tk = Ticket("sparse", problem, order, K)
        with self.cv:
            failed = self.sparse_failed
            if failed is None:
                self.sparse_q.append(tk)
                self.cv.notify_all()
        if failed is not None:                  # the engine thread stopped: answered at once with its error
            self._finish([tk], error=failed)
        return tk

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for th in self.planner_threads + self.threads:
            th.join()
        self.engine_thread.join()               # may still hand requests to the single-query lane
        with self.cv:
            self.single_closed = True
            self.cv.notify_all()
        self.single_thread.join()
        self.sparse_engine.close()
        # the blocks cached for this server's streams are of no use to later streams: give them back
        t = _lib.torch()
        t.cuda.synchronize()
        t.cuda.empty_cache()

    # ------------------------------------------------------------------ lanes
    def _finish(self, tickets, results=None, error=None):
        now = time.perf_counter()
        with self.cv:
            for k, tk in enumerate(tickets):
                tk.t_done = now
                if error is not None:
                    tk.error = error
                else:
                    tk.result = results[k]
                self.in_use -= tk.footprint
                tk._event.set()
                self.done.append(tk)
            self.cv.notify_all()

    @staticmethod
    def _plan(tk):
        if tk.plans is None:
            order = tk.order if tk.order is not None else auto.planning_order(tk.problem)
            tk.order = order
            orders, flat, owner = batch.plan_problems([tk.problem], [order])
            tk.plans = (flat, owner)
            tk.footprint = _dense_footprint([p for p in flat if p.feasible])
        return tk

    PLAN_CHUNK = 8                              # requests a planner takes at once

    def _planner(self):
        """Plans waiting dense requests a few at a time and hands them to the lanes (dense_q)."""
        while True:
            with self.cv:
                while not self.plan_q and not self.closed:
                    self.cv.wait()
                if not self.plan_q:
                    return
                take = [self.plan_q.popleft() for _ in range(min(self.PLAN_CHUNK, len(self.plan_q)))]
                self.planning += len(take)
            planned = []
            for tk in take:
                try:
                    planned.append(self._plan(tk))
                except Exception as exc:          # this request alone is answered with its error
                    tk.plans, tk.footprint = None, 0
                    self._finish([tk], error=exc)
            with self.cv:
                self.planning -= len(take)
                self.dense_q.extend(planned)
                self.cv.notify_all()

    def _dense_lane(self, stream, wide_stream):
        t = _lib.torch()
        while True:
            with self.cv:
                self.idle_lanes += 1
                while not self.dense_q and not (self.closed and not self.plan_q and not self.planning):
                    self.cv.wait()
                self.idle_lanes -= 1
                if not self.dense_q:
                    return                        # closed, and nothing left to plan
                n = claim_count(len(self.dense_q), self.idle_lanes, self.max_batch, self.share_above)
                claimed = [self.dense_q.popleft() for _ in range(n)]
            while claimed:
                with self.cv:
                    # admit the longest prefix that fits the free budget (at least one request when idle)
                    take, need = [], 0
                    for tk in claimed:
                        if self.in_use + need + tk.footprint > self.budget and (take or self.in_use > 0):
                            break
                        take.append(tk)
                        need += tk.footprint
                    if not take:
                        self.cv.wait(0.005)
                        continue
                    self.in_use += need
                    now = time.perf_counter()
                    for tk in take:
                        tk.t_start = now
                claimed = claimed[len(take):]
                probs = [tk.problem for tk in take]
                orders = [tk.order for tk in take]
                flat, owner = [], []
                for k, tk in enumerate(take):
                    f, o = tk.plans
                    flat += f
                    owner += [(k, i) for (_, i) in o]
                try:
                    with t.cuda.stream(stream):
                        recs = batch.certify_planned(probs, orders, flat, owner, stream=stream,
                                                     wide_stream=wide_stream)
                    self._finish(take, recs)
                except Exception as exc:          # delivered to every waiter of the failed batch
                    self._finish(take, error=exc)

    def _sparse_engine_loop(self, stream):
        """Admits waiting sparse requests while the engine's arena is less than SPARSE_ADMIT full (always one when
        nothing runs), steps the engine and completes finished requests.  A request that finds the arena full while
        others run waits at the head of the queue for memory to free; one that finds it full alone is answered by
        the checkpointed one-at-a-time sweeps."""
        eng = self.sparse_engine
        owner, waiting = {}, collections.deque()
        running = set()

        def complete(r):
            if r.error is not None:
                self._finish([r.tag], error=r.error)
            else:
                self._finish([r.tag], [r.record])

        try:
            self._engine_rounds(eng, owner, waiting, running, complete)
        except Exception as exc:                # the engine itself failed: nothing more runs on it
            with self.cv:
                self.sparse_failed = exc
                queued = list(self.sparse_q)
                self.sparse_q.clear()
            for tk in queued:
                self._finish([tk], error=exc)
            for r in list(owner.values()) + list(waiting) + list(running):
                if r.tag.t_done is None:
                    self._finish([r.tag], error=exc)
        finally:
            for r in list(owner.values()) + list(waiting):
                if r.tag.t_done is None:
                    self._finish([r.tag], error=RuntimeError("server closed"))

    def _engine_rounds(self, eng, owner, waiting, running, complete):
        """The engine thread's loop (see _sparse_engine_loop); returns when the server closes."""
        while True:
            with self.cv:
                while not self.sparse_q and not owner and not waiting and not self.closed:
                    self.cv.wait()
                if self.closed and not self.sparse_q and not owner and not waiting:
                    return
                new = list(self.sparse_q)
                self.sparse_q.clear()
            now = time.perf_counter()
            for tk in new:                      # host planning (C++ trellis words) as the requests arrive
                tk.t_start = now
                r = sparse_batch.Request(tk.problem, tk.K, tk.order, tag=tk, deepen=self.sparse_deepen)
                try:
                    try:
                        r.plan()
                    except ValueError as exc:
                        if not wide_trellis(exc):
                            raise
                        self._to_single(r)          # beyond the device states' row slots: the search answers it
                        continue
                    if len(r.trs) > eng.max_jobs:   # more queries than the engine has slots: one at a time
                        self._to_single(r)
                        continue
                    elif r.trs:
                        waiting.append(r)
                        continue
                    else:
                        r.finish(eng)
                except Exception as exc:
                    r.error = exc
                complete(r)
            st = eng.stats()
            while waiting and (not running or (st["in_use"] < self.sparse_admit * st["arena_bytes"] and
                                               not st["nomem_last"] and
                                               len(owner) + len(waiting[0].trs) <= eng.max_jobs)):
                r = waiting.popleft()
                try:
                    r.admit(eng, owner)
                except Exception as exc:        # answered with the error; queries already added drain first
                    r.error = exc
                    if not r.ids:
                        complete(r)
                        continue
                running.add(r)
                st = eng.stats()
            if not owner:
                continue
            try:
                eng.step()                      # on the engine's own stream
            except Exception as exc:            # a failed step ends every running request with its error
                for r in list(running):
                    for rid in r.ids:
                        owner.pop(rid, None)
                    r.error = exc
                    complete(r)
                running.clear()
                continue
            for rid in eng.poll():
                r = owner.pop(rid, None)
                if r is None:                   # a query of a request already answered with an error
                    eng.release(rid)
                    continue
                r.collect(eng, rid)
                r.left -= 1
                if r.left:
                    continue
                running.discard(r)
                alone = not running
                r.finish(eng)
                if r.pending:                   # no solution at this threshold: admitted again for the next
                    waiting.appendleft(r)
                    continue
                if isinstance(r.error, MemoryError):
                    r.error = None
                    if not alone:
                        waiting.appendleft(r)   # again when memory has freed
                        continue
                    self._to_single(r)          # too large for the arena alone: checkpointed sweeps
                    continue
                complete(r)

    def _to_single(self, r):
        with self.cv:
            self.single_q.append(r)
            self.cv.notify_all()

    def _sparse_single_lane(self):
        """Answers the requests handed over by the engine thread with sparse_certify (checkpointed single-query
        sweeps at the server's state budget), one at a time on this lane's stream.  The stream is made at the first
        request: most mixes never need it, and every stream beyond the device's hardware queues shares one."""
        t = _lib.torch()
        stream = None
        while True:
            with self.cv:
                while not self.single_q and not self.single_closed:
                    self.cv.wait()
                if not self.single_q:
                    return
                r = self.single_q.popleft()
            try:
                if stream is None:
                    stream = t.cuda.Stream()
                with t.cuda.stream(stream):
                    rec = sparse_certify(r.problem, r.K, r.order, max_states=self.sparse_max_states)
                    stream.synchronize()
                self._finish([r.tag], [rec])
            except Exception as exc:
                self._finish([r.tag], error=exc)

    # ------------------------------------------------------------------ statistics
    def latencies(self, kind=None):
        return [tk.latency for tk in self.done if tk.latency is not None and (kind is None or tk.kind == kind)]


def claim_count(waiting, idle_lanes, max_batch, share_above=64):
    """Requests a waking dense lane takes from a queue of `waiting`: a long queue (share_above or more) is shared
    with the lanes that are idle too, so that they plan and run in parallel; a short one (trickling arrivals) is
    taken whole; never more than max_batch."""
    n = waiting
    if n >= share_above:
        n = -(-n // (1 + idle_lanes))
    return min(max_batch, n)


def wide_trellis(exc):
    """Is this the error of a trellis with more row slots than a device state holds?"""
    return isinstance(exc, ValueError) and "bits per state" in str(exc)


def search_certify(p, K, order=None, k_table=2):
    """The record of sparse_certify for a problem whose trellis is too wide for the sparse states: the least weight
    <= K of a solution over all observables by the weight-bounded search on the GPU (search.distance, thresholds 1..K
    on label-free tables, the gm_search route of the reference's pipeline for wide codes).  The observable is the
    lowest one the witness flips."""
    from . import search
    from .problem import verify_witness
    r = search.distance(p, k_table=k_table, kmax=K, start=1, labelfree=True)
    if r["distance"] is None or not r["solutions"]:
        return {"distance": None, "witness": None, "observable": None, "witness_ok": True, "K": K, "path": "search",
                "peak_states": 0}
    cols = sorted(r["solutions"][0])
    lg = 0
    for j in cols:
        lg ^= p.l[j]
    obs = (lg & -lg).bit_length() - 1
    v = r["distance"]
    return {"distance": v, "witness": cols, "observable": obs, "K": K,
            "witness_ok": verify_witness(p, cols, v, obs), "path": "search", "peak_states": 0}


def sparse_certify(p, K, order=None, max_states=30_000_000, deepen=False):
    """Least weight <= K of a solution over all observables by the sparse MITM DP: {"distance", "witness",
    "observable", "witness_ok"}.  One query at a time on the single-query path: the observables are swept in turn,
    after a solution of weight v the others at v - 1 (auto.sparse_distance's loop, cdist/auto.py:257-284, with the
    default split).  deepen: first through the lower thresholds of sparse_batch.threshold_ladder(K), as the
    engine's requests do (the same result; one query at a time it is slower on the serving mix, whose codes mostly
    have distances above K).  A table beyond max_states raises sparse_dp.BudgetExceeded (the request is answered
    with that error).  A problem whose trellis needs more row slots than a device state holds is answered by the
    weight-bounded search instead (search_certify, path "search")."""
    from .ordering import greedy_order_native
    from .problem import verify_witness
    order = order if order is not None else greedy_order_native(p.h, p.nrows, p.words())
    try:
        trs = [sparse_dp.Trellis.for_query(p, i, order) for i in range(p.nobs)]
    except ValueError as exc:
        if not wide_trellis(exc):
            raise
        return search_certify(p, K, order)
    best, peak = None, 0
    for kt in (sparse_batch.threshold_ladder(K) if deepen else [K]):
        for i, tr in enumerate(trs):
            if not tr.feasible:
                continue
            k = kt if best is None else best[0] - 1
            if k < 1:
                break
            r = sparse_dp.query(tr, k, mode="mitm", use_h=False, witness=True, backend="hip", max_states=max_states)
            peak = max(peak, r["peak_states"])
            if r["value"] is not None:
                best = (r["value"], r["witness"], i)
        if best is not None:
            break
    if best is None:
        return {"distance": None, "witness": None, "observable": None, "witness_ok": True, "K": K, "path": "single",
                "peak_states": peak}
    v, cols, obs = best
    return {"distance": v, "witness": cols, "observable": obs, "K": K,
            "witness_ok": verify_witness(p, cols, v, obs), "path": "single", "peak_states": peak}


def percentile(xs, q):
    if not xs:
        return None
    xs = sorted(xs)
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def run_poisson(server, requests, rate, seed=0, progress=None):
    """Submit requests at Poisson arrival times of the given rate (per second); wait for all; returns
    (tickets, wall seconds).  progress(msg) reports every 1000th submission and, while waiting, answered requests."""
    import numpy as np
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(requests))
    tickets = []
    t0 = time.perf_counter()
    at = t0
    for (kind, args), gap in zip(requests, gaps):
        at += gap
        now = time.perf_counter()
        if at > now:
            time.sleep(at - now)
        if kind == "dense":
            tickets.append(server.submit_dense(*args))
        else:
            tickets.append(server.submit_sparse(*args))
        if progress is not None and len(tickets) % 1000 == 0:
            progress("%d submitted, %d answered" % (len(tickets), sum(tk._event.is_set() for tk in tickets)))
    wait_all(tickets, progress)
    return tickets, time.perf_counter() - t0


def wait_all(tickets, progress=None, every=20.0):
    """Wait until every ticket is answered.  progress(msg) is called about every `every` seconds, and only when more
    tickets have been answered since its last call."""
    i, last, t_last = 0, -1, time.perf_counter()
    while i < len(tickets):
        if tickets[i]._event.wait(timeout=1.0):
            i += 1
        if progress is not None and time.perf_counter() - t_last >= every:
            t_last = time.perf_counter()
            done = sum(tk._event.is_set() for tk in tickets)
            if done > last:
                per = {}
                for tk in tickets:
                    a, n = per.get(tk.kind, (0, 0))
                    per[tk.kind] = (a + tk._event.is_set(), n + 1)
                progress("%d of %d answered (%s)" % (done, len(tickets), ", ".join(
                    "%s %d/%d" % (k, a, n) for k, (a, n) in sorted(per.items()))))
                last = done
