"""Host logic of serving without a device: the threshold ladder of sparse requests, a request's deepening from one
threshold to the next (against a stand-in engine that answers each query with its least weight when it is within
the threshold), and the waiting helper's progress reports."""
import threading
import time

from tropidist_amd import problem, serve, sparse_batch


def test_threshold_ladder():
    assert sparse_batch.threshold_ladder(4) == [4]
    assert sparse_batch.threshold_ladder(3) == [3]
    assert sparse_batch.threshold_ladder(5) == [5]
    assert sparse_batch.threshold_ladder(7) == [5, 7]
    assert sparse_batch.threshold_ladder(8) == [4, 6, 8]
    assert sparse_batch.threshold_ladder(10) == [4, 6, 8, 10]
    for K in range(1, 20):
        ks = sparse_batch.threshold_ladder(K)
        assert ks[-1] == K and ks == sorted(set(ks)) and (len(ks) == 1 or ks[0] >= 4)


class _Engine:
    """Answers a query with value `least[obs]` when it is <= the query's threshold, else with no value."""

    def __init__(self, least, witness):
        self.least, self.wit = least, witness
        self.added, self.released, self.q = [], [], {}

    def add_words(self, prm, tm, N, K):
        rid = len(self.added)
        self.added.append(K)
        self.q[rid] = (prm, K)
        return rid

    def result(self, rid):
        obs, K = self.q[rid]
        v = self.least[obs]
        return sparse_batch.DONE, (v if v is not None and v <= K else None), {}

    def witness(self, rid, order, value):
        return self.wit

    def release(self, rid):
        self.released.append(rid)


def _run(req, eng):
    """Admits and finishes the request until it is no longer pending; returns the thresholds swept."""
    while True:
        req.admit(eng)
        for rid in req.ids:                       # as the engine's poll hands them back
            req.collect(eng, rid)
        req.finish(eng)
        if not req.pending:
            return eng.added


def test_request_deepens_until_a_threshold_finds_a_solution():
    p = problem.chain(6)                          # least weight 6: every column
    witness = list(range(p.ncols))
    # one query per observable; the stand-in takes the observable as the query's words
    req = sparse_batch.Request(p, 8, order=list(range(p.ncols)))
    req.trs = [(0, 0, None)]
    eng = _Engine({0: 6}, witness)
    assert _run(req, eng) == [4, 6]               # 4 finds nothing, 6 finds the solution; 8 is never swept
    assert req.record["distance"] == 6 and req.record["observable"] == 0 and req.record["witness_ok"]
    assert eng.released == [0, 1]                 # the sweep at 4 as soon as it finished, the one at 6 after its witness


def test_request_without_a_solution_sweeps_every_threshold():
    p = problem.chain(6)
    req = sparse_batch.Request(p, 10, order=list(range(p.ncols)))
    req.trs = [(0, 0, None)]
    eng = _Engine({0: None}, None)
    assert _run(req, eng) == [4, 6, 8, 10]
    assert req.record["distance"] is None and req.record["K"] == 10


def test_request_picks_the_first_observable_with_the_least_value():
    p = problem.chain(6)
    req = sparse_batch.Request(p, 8, order=list(range(p.ncols)))
    req.trs = [(0, 0, None), (1, 1, None), (2, 2, None)]
    eng = _Engine({0: 7, 1: 6, 2: 6}, list(range(p.ncols)))
    assert _run(req, eng) == [4, 4, 4, 6, 6, 6]
    assert req.record["distance"] == 6 and req.record["observable"] == 1


def test_wait_all_reports_only_progress():
    class T:
        def __init__(self, kind):
            self.kind, self._event = kind, threading.Event()

    tickets = [T("dense" if i % 3 else "sparse") for i in range(30)]
    msgs = []

    def answer():
        for tk in reversed(tickets):
            time.sleep(0.02)
            tk._event.set()

    th = threading.Thread(target=answer)
    th.start()
    serve.wait_all(tickets, msgs.append, every=0.1)
    th.join()
    assert all(tk._event.is_set() for tk in tickets)
    assert msgs and all("answered" in m for m in msgs)
    counts = [int(m.split()[0]) for m in msgs]
    assert counts == sorted(set(counts))          # strictly increasing: no report without new answers


def test_claim_count_shares_long_queues_only():
    assert serve.claim_count(10, 2, 1024) == 10            # short queue: taken whole
    assert serve.claim_count(540, 2, 1024) == 180          # long queue: a third with two idle lanes
    assert serve.claim_count(540, 0, 1024) == 540          # no idle lane: all of it
    assert serve.claim_count(5000, 0, 1024) == 1024        # never more than max_batch
    assert serve.claim_count(64, 1, 1024, share_above=64) == 32
    assert serve.claim_count(63, 1, 1024, share_above=64) == 63


def test_a_query_out_of_memory_cancels_the_others_and_fails_the_request():
    p = problem.chain(6)
    req = sparse_batch.Request(p, 4, order=list(range(p.ncols)))
    req.trs = [(0, 0, None), (1, 1, None), (2, 2, None)]

    class Eng(_Engine):
        def __init__(self):
            super().__init__({0: 4, 1: None, 2: 4}, list(range(p.ncols)))
            self.cancelled = []

        def result(self, rid):
            if rid == 1:
                return sparse_batch.NOMEM, None, {}
            return super().result(rid)

        def cancel(self, rid):
            self.cancelled.append(rid)

    eng = Eng()
    req.admit(eng)
    req.collect(eng, 1)                           # the out-of-memory query is polled first
    assert eng.cancelled == [0, 2] and eng.released == [1]
    for rid in (0, 2):
        req.collect(eng, rid)
    req.finish(eng)
    assert isinstance(req.error, MemoryError) and req.record is None and not req.pending
    assert sorted(eng.released) == [0, 1, 2]


class _BrokenEngine:
    """A stand-in engine whose step raises: the engine thread must answer every sparse ticket with the error (the
    running ones, the queued ones and any submitted afterwards) and never leave a waiter hanging."""
    max_jobs = 64
    max_states = None

    def stats(self):
        return {"in_use": 0, "arena_bytes": 1 << 30, "nomem_last": 0}

    def add_words(self, prm, tm, N, K):
        self.n += 1
        return self.n

    def __init__(self, fail_step):
        self.n, self.fail_step = 0, fail_step

    def step(self):
        if self.fail_step:
            raise RuntimeError("device lost")
        return 1

    def poll(self):
        raise RuntimeError("device lost")

    def cancel(self, rid):
        pass

    def release(self, rid):
        pass


def _bare_server(engine):
    """A Server with its queues and engine thread only (no device, no dense lanes)."""
    import collections
    srv = object.__new__(serve.Server)
    srv.cv = threading.Condition()
    srv.sparse_q, srv.dense_q, srv.single_q = collections.deque(), collections.deque(), collections.deque()
    srv.sparse_failed, srv.closed, srv.done, srv.in_use = None, False, [], 0
    srv.sparse_engine, srv.sparse_admit, srv.sparse_deepen = engine, 0.5, False
    return srv


def _expect_error(tk, timeout):
    try:
        tk.wait(timeout=timeout)
    except RuntimeError as exc:
        assert "device lost" in str(exc)
    else:
        raise AssertionError("a request on a failed engine was answered with a record")


def test_engine_thread_failure_answers_every_sparse_ticket():
    """A failure outside the step (here the poll) stops the engine thread: every ticket gets the error."""
    srv = _bare_server(_BrokenEngine(fail_step=False))
    p = problem.quasi_cyclic(3, 3, seed=1)
    first = [srv.submit_sparse(p, 4) for _ in range(3)]
    th = threading.Thread(target=srv._sparse_engine_loop, args=(None,), daemon=True)
    th.start()
    for tk in first:
        _expect_error(tk, 30)
    th.join(timeout=30)
    assert not th.is_alive() and isinstance(srv.sparse_failed, RuntimeError)
    _expect_error(srv.submit_sparse(p, 4), 1)     # after the failure: answered at once


def test_failed_step_answers_running_requests_and_keeps_serving():
    srv = _bare_server(_BrokenEngine(fail_step=True))
    p = problem.quasi_cyclic(3, 3, seed=1)
    th = threading.Thread(target=srv._sparse_engine_loop, args=(None,), daemon=True)
    th.start()
    for tk in [srv.submit_sparse(p, 4) for _ in range(3)]:
        _expect_error(tk, 30)
    _expect_error(srv.submit_sparse(p, 4), 30)    # a later request is admitted and answered too
    with srv.cv:
        srv.closed = True
        srv.cv.notify_all()
    th.join(timeout=30)
    assert not th.is_alive() and srv.sparse_failed is None


def test_ticket_wait_times_out():
    tk = serve.Ticket("dense", None, None, None)
    t0 = time.time()
    try:
        tk.wait(timeout=0.05)
    except TimeoutError:
        assert time.time() - t0 < 5
    else:
        raise AssertionError("wait returned without an answer")
