"""GPU parity of the batched sparse engine (many meet-in-the-middle queries per launch, tdp_spb_*): values equal
brute force and the one-query device path on random problems; each result is the same alone and in a mix whose
queries are admitted while others run; witnesses check against H and L; whole requests equal serve.sparse_certify;
the state budget and a full arena behave as documented."""
import random

import pytest

from sparse_helpers import brute, check_witness, random_problem
from tropidist_amd import problem, serve, sparse_batch, sparse_dp

pytestmark = pytest.mark.gpu


def _drain(eng):
    steps = 0
    while eng.step():
        steps += 1
        assert steps < 100000
    return steps


def _random_cases(seed, nproblems):
    rng = random.Random(seed)
    cases = []
    for _ in range(nproblems):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            if not tr.feasible:
                continue
            truth = brute(p, obs)
            for K in range(1, p.ncols + 2):
                cases.append((p, obs, tr, K, truth if truth is not None and truth <= K else None))
    return cases


def _alone(tr, K):
    eng = sparse_batch.Engine(arena_bytes=32 << 20, max_jobs=4)
    try:
        rid = eng.add(tr, K)
        _drain(eng)
        st, v, _ = eng.result(rid)
        assert st == sparse_batch.DONE
        w = eng.witness(rid, tr, v) if v is not None else None
        eng.release(rid)
        return v, w
    finally:
        eng.close()


def test_random_queries_in_one_batch_against_brute_force():
    print("case sparse batch random seed 5", flush=True)
    cases = _random_cases(5, 50)
    assert len(cases) > 300
    eng = sparse_batch.Engine(arena_bytes=512 << 20)
    try:
        ids = [eng.add(tr, K) for _, _, tr, K, _ in cases]
        _drain(eng)
        for (p, obs, tr, K, want), rid in zip(cases, ids):
            st, v, _ = eng.result(rid)
            assert st == sparse_batch.DONE and v == want, (K, v, want)
            if want is not None:
                check_witness(p, eng.witness(rid, tr, v), want, obs)
            eng.release(rid)
    finally:
        eng.close()


def test_same_as_one_query_path():
    print("case sparse batch vs sparse_dp.query seed 6", flush=True)
    cases = _random_cases(6, 12)
    eng = sparse_batch.Engine(arena_bytes=256 << 20)
    try:
        ids = [eng.add(tr, K) for _, _, tr, K, _ in cases]
        _drain(eng)
        for (p, obs, tr, K, _), rid in zip(cases, ids):
            r = sparse_dp.query(tr, K, mode="mitm", use_h=False, backend="hip")
            assert eng.result(rid)[1] == r["value"]
            eng.release(rid)
    finally:
        eng.close()


def test_each_result_same_alone_and_in_a_mix_admitted_mid_run():
    """Queries join the engine in waves while earlier ones are mid-sweep; every value and witness equals the
    query's result when it runs alone."""
    print("case sparse batch continuous admission seed 7", flush=True)
    cases = _random_cases(7, 10)
    qc = problem.quasi_cyclic(6, 6, seed=3001)
    from tropidist_amd.ordering import greedy_straddle_order
    order = greedy_straddle_order(qc.h, qc.nrows)
    for obs in range(qc.nobs):
        tr = sparse_dp.Trellis.for_query(qc, obs, order)
        if tr.feasible:
            cases.append((qc, obs, tr, 6, None))
    rng = random.Random(7)
    rng.shuffle(cases)
    alone = [_alone(tr, K) for _, _, tr, K, _ in cases]
    eng = sparse_batch.Engine(arena_bytes=512 << 20)
    try:
        ids, k = [None] * len(cases), 0
        while k < len(cases) or eng.active:
            for _ in range(rng.randint(0, 12)):
                if k < len(cases):
                    ids[k] = eng.add(cases[k][2], cases[k][3])
                    k += 1
            eng.step()
        for (p, obs, tr, K, _), rid, (v0, w0) in zip(cases, ids, alone):
            st, v, _ = eng.result(rid)
            assert st == sparse_batch.DONE and v == v0
            if v is not None:
                w = eng.witness(rid, tr, v)
                assert w == w0
                check_witness(p, w, v, obs)
            eng.release(rid)
    finally:
        eng.close()


def test_requests_equal_sparse_certify():
    print("case sparse batch requests on QC codes", flush=True)
    reqs = [(problem.quasi_cyclic(l, m, seed=s), K) for l, m, s, K in
            [(6, 6, 3001, 6), (6, 6, 3002, 8), (6, 8, 3003, 8), (9, 8, 3004, 6), (8, 9, 3005, 6), (6, 9, 3006, 7)]]
    got = sparse_batch.certify_requests(reqs, arena_bytes=1 << 30)
    for (p, K), r in zip(reqs, got):
        assert r.error is None, r.error
        want = serve.sparse_certify(p, K)
        assert r.record["distance"] == want["distance"]
        assert r.record["observable"] == want["observable"]
        assert r.record["witness_ok"]
        deep = serve.sparse_certify(p, K, deepen=True)
        assert deep["distance"] == want["distance"] and deep["observable"] == want["observable"]
        # the engine deepens through the threshold ladder: the least value equals one sweep of every observable at K
        vals = [sparse_dp.query(tr, K, mode="mitm", use_h=False, backend="hip")["value"]
                for tr in (sparse_dp.Trellis.for_query(p, i, r.order) for i in range(p.nobs)) if tr.feasible]
        vals = [v for v in vals if v is not None]
        assert r.record["distance"] == (min(vals) if vals else None)


def test_small_arena_queries_wait_for_memory():
    print("case sparse batch requests in an arena of 6 MiB", flush=True)
    reqs = [(problem.quasi_cyclic(l, m, seed=s), K) for l, m, s, K in
            [(6, 6, 3001, 6), (6, 8, 3003, 8), (9, 8, 3004, 6), (8, 9, 3005, 6), (6, 9, 3006, 7), (9, 9, 3007, 8)]]
    eng = sparse_batch.Engine(arena_bytes=6 << 20)
    try:
        got = sparse_batch.certify_requests(reqs, engine=eng)
        st = eng.stats()
        print("arena peak %d of %d bytes, %d queries gave their memory up" % (st["peak_in_use"], st["arena_bytes"],
                                                                            st["nomem"]), flush=True)
        assert st["running"] == 0 and st["in_use"] == 0
    finally:
        eng.close()
    for (p, K), r in zip(reqs, got):
        assert r.error is None, r.error
        want = serve.sparse_certify(p, K)
        assert r.record["distance"] == want["distance"] and r.record["observable"] == want["observable"]
        assert r.record["witness_ok"]


def test_state_budget_raises_like_one_query_path():
    print("case sparse batch state budget", flush=True)
    p = problem.quasi_cyclic(6, 6, seed=3001)
    from tropidist_amd.ordering import greedy_straddle_order
    order = greedy_straddle_order(p.h, p.nrows)
    tr = next(t for t in (sparse_dp.Trellis.for_query(p, i, order) for i in range(p.nobs)) if t.feasible)
    with pytest.raises(sparse_dp.BudgetExceeded):
        sparse_dp.query(tr, 6, mode="mitm", use_h=False, backend="hip", max_states=20)
    eng = sparse_batch.Engine(arena_bytes=64 << 20, max_states=20)
    try:
        got = sparse_batch.certify_requests([(p, 6, order)], engine=eng)
        assert isinstance(got[0].error, sparse_dp.BudgetExceeded)
    finally:
        eng.close()


def test_full_arena_falls_back_to_one_query_path():
    print("case sparse batch arena of 1 MiB", flush=True)
    p = problem.quasi_cyclic(9, 8, seed=3004)
    eng = sparse_batch.Engine(arena_bytes=1 << 20)
    try:
        got = sparse_batch.certify_requests([(p, 7)], engine=eng)
        assert got[0].error is None
        want = serve.sparse_certify(p, 7)
        assert got[0].record["distance"] == want["distance"]
        eng.step()
        assert eng.active == 0
    finally:
        eng.close()
