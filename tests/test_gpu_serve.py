"""Serving loop on the GPU: every request's result is the same alone and in the mix (continuous batching,
overlapping lanes, footprint admission), and a short Poisson stream completes with recorded latencies."""
import numpy as np
import pytest

from tropidist_amd import auto, problem, serve, workloads
from tropidist_amd.problem import verify_witness

pytestmark = pytest.mark.gpu


def _mix(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if i % 7 == 3:
            p = problem.quasi_cyclic(int(rng.choice([3, 6, 7])), int(rng.choice([3, 6, 7])), seed=seed + i)
            out.append(("sparse", p, int(rng.integers(3, 7))))
        else:
            out.append(("dense", workloads.banded_instance(rng, seed + i, int(rng.integers(6, 15)), bool(i % 2)), None))
    return out


def _alone(kind, p, K):
    if kind == "dense":
        r = auto.certify(p)
        return r["distance"], r["per_observable"]
    r = serve.sparse_certify(p, K)
    return r["distance"], None


def test_results_alone_equal_results_in_the_mix():
    print("case serve mix", flush=True)
    reqs = _mix(42, 9100)
    srv = serve.Server(memory_bytes=8 << 30, dense_lanes=2, max_batch=8, sparse_arena_bytes=1 << 30)
    tickets = [srv.submit_dense(p) if kind == "dense" else srv.submit_sparse(p, K) for kind, p, K in reqs]
    recs = [tk.wait(timeout=300) for tk in tickets]
    srv.close()
    for (kind, p, K), rec in zip(reqs, recs):
        d, per = _alone(kind, p, K)
        assert rec["distance"] == d, (kind, p.name)
        if kind == "dense":
            assert rec["per_observable"] == per
        if d is not None:
            assert rec["witness_ok"] and verify_witness(p, rec["witness"], d, rec["observable"])


def test_poisson_stream_completes_with_latencies():
    print("case serve poisson", flush=True)
    reqs = [("dense", (p,)) if kind == "dense" else ("sparse", (p, K)) for kind, p, K in _mix(30, 9300)]
    srv = serve.Server(memory_bytes=8 << 30, dense_lanes=2, sparse_arena_bytes=1 << 30)
    tickets, wall = serve.run_poisson(srv, reqs, rate=50.0, seed=1)
    srv.close()
    lat = [tk.latency for tk in tickets]
    assert all(x is not None and x > 0 for x in lat) and all(tk.error is None for tk in tickets)
    assert serve.percentile(lat, 50) <= serve.percentile(lat, 99)


def test_sparse_requests_beyond_the_engine_slots():
    """Twelve sparse requests whose queries far exceed an engine of 6 job slots: admission waits for slots, and
    every result equals the one-at-a-time answer."""
    print("case serve sparse slots", flush=True)
    probs = [(problem.quasi_cyclic(l, m, seed=9500 + k), K) for k, (l, m, K) in
             enumerate([(6, 6, 6), (6, 8, 6), (6, 6, 7), (9, 6, 5), (6, 9, 6), (8, 6, 6)] * 2)]
    srv = serve.Server(memory_bytes=8 << 30, dense_lanes=1, sparse_arena_bytes=1 << 30, sparse_max_jobs=6)
    tickets = [srv.submit_sparse(p, K) for p, K in probs]
    recs = [tk.wait(timeout=300) for tk in tickets]
    srv.close()
    for (p, K), rec in zip(probs, recs):
        want = serve.sparse_certify(p, K)
        assert rec["distance"] == want["distance"] and rec["observable"] == want["observable"]
        if rec["distance"] is not None:
            assert rec["witness_ok"]


def test_unplannable_dense_request_fails_alone():
    """A dense request whose plan needs more than 63 active rows (a dense random 80 x 200 check matrix) is answered
    with the planner's ValueError; the good requests claimed with it are answered normally."""
    import random
    print("case serve unplannable request", flush=True)
    rng = random.Random(5)
    bad = problem.Problem("wide", 80, 1, [rng.getrandbits(80) | 1 for _ in range(200)],
                          [rng.getrandbits(1) for _ in range(200)])
    good = [workloads.banded_instance(np.random.default_rng(s), 9300 + s, 10, s % 2 == 1) for s in range(6)]
    srv = serve.Server(memory_bytes=8 << 30, dense_lanes=1, max_batch=16, sparse_arena_bytes=1 << 30)
    try:
        tickets = [srv.submit_dense(p) for p in good[:3]] + [srv.submit_dense(bad)] + \
                  [srv.submit_dense(p) for p in good[3:]]
        for tk, p in zip(tickets[:3] + tickets[4:], good):
            rec = tk.wait(timeout=300)
            assert rec["distance"] == auto.certify(p)["distance"] and rec["witness_ok"]
        with pytest.raises(ValueError):
            tickets[3].wait(timeout=300)
    finally:
        srv.close()


def _cfg5_requests(n, seed=5000):
    """The first n requests of the serving configuration's mix (dense widths 10-25, sparse MITM with K 4-10 on
    quasi-cyclic codes with l, m in 6-15), built as the benchmark builds them."""
    out = []
    for kind, a in workloads.serving_mix(n, seed=seed, K=(4, 10)):
        if kind == "sparse":
            out.append(("sparse", problem.quasi_cyclic(a["l"], a["m"], seed=a["seed"]), a["K"]))
        else:
            rng = np.random.default_rng(a["seed"])
            out.append(("dense", workloads.banded_instance(rng, a["seed"], a["width"], a["weighted"]), None))
    return out


def test_serving_configuration_shape_alone_equals_mix_and_oracle():
    """Requests of the serving configuration itself: served together, each equals its one-at-a-time answer (the
    same error where a sparse table passes the state budget), every witness checks against H and L, and the
    oracle's CPU kernels agree where they run in seconds: dense plans of work <= 2e8 (or.run_ops, every
    observable), sparse requests of at most 56 row slots and 150 columns (the oracle's tables, the same sweeps)."""
    import oracle
    from tropidist_amd import ordering, sparse_dp, trellis
    print("case serve cfg5 shape", flush=True)
    reqs = _cfg5_requests(90)
    assert {k for k, _, _ in reqs} == {"dense", "sparse"}
    states = 4_000_000
    srv = serve.Server(memory_bytes=48 << 30, dense_lanes=3, sparse_arena_bytes=16 << 30, sparse_max_states=states)
    try:
        tickets = [srv.submit_dense(p) if kind == "dense" else srv.submit_sparse(p, K) for kind, p, K in reqs]
        for tk in tickets:
            assert tk._event.wait(600), "request not answered"
    finally:
        srv.close()
    checked = {"dense_oracle": 0, "sparse_oracle": 0}
    for (kind, p, K), tk in zip(reqs, tickets):
        if kind == "dense":
            alone = auto.certify(p)
            assert tk.error is None, tk.error
            rec = tk.result
            assert rec["distance"] == alone["distance"] and rec["per_observable"] == alone["per_observable"]
            if rec["distance"] is not None:
                assert rec["witness_ok"] and verify_witness(p, rec["witness"], rec["distance"], rec["observable"])
            order = auto.planning_order(p)
            plans = trellis.build_plans_native(p.h, p.l, [1 << i for i in range(p.nobs)], order, p.nrows, p.w)
            if all(pl.work() <= 2e8 for pl in plans if pl.feasible):
                want = [oracle.run_plan_c(pl, "oracle") if pl.feasible else None for pl in plans]
                assert want == rec["per_observable"]
                checked["dense_oracle"] += 1
            continue
        try:
            alone = serve.sparse_certify(p, K, max_states=states)
        except (sparse_dp.BudgetExceeded, ValueError) as exc:
            # both paths refuse the request: a table past the state budget, or more row slots than a state holds
            assert type(tk.error) is type(exc), (tk.error, exc)
            continue
        assert tk.error is None, tk.error
        assert tk.result["distance"] == alone["distance"]
        if alone["distance"] is not None:
            assert tk.result["witness_ok"] and alone["witness_ok"]
        if alone["path"] == "search":            # a trellis beyond the device states: no sparse tables to compare
            continue
        order = ordering.greedy_order_native(p.h, p.nrows, p.words())
        trs = [sparse_dp.Trellis.for_query(p, i, order) for i in range(p.nobs)]
        if p.ncols <= 150 and all(t.nslots <= 56 for t in trs if t.feasible):
            best = None
            for tr in trs:
                if not tr.feasible:
                    continue
                k = K if best is None else best - 1
                if k < 1:
                    break
                r = sparse_dp.query(tr, k, mode="mitm", use_h=False, witness=False, backend=oracle.OracleTables(tr))
                if r["value"] is not None:
                    best = r["value"]
            assert best == alone["distance"]
            checked["sparse_oracle"] += 1
    print("  oracle-checked", checked, flush=True)
    assert checked["dense_oracle"] >= 30


WIDE_REQUESTS = [(13, 15, 5076, 5), (15, 11, 6440, 9), (15, 15, 6809, 10)]   # trellises of 122-130 row slots


def test_requests_beyond_the_state_width_go_to_the_search():
    """Sparse requests of the serving mix whose trellises need more row slots than a device state holds (122-130 >
    120) are answered by the weight-bounded search on the GPU (gm_search's route), alone and through the server
    alike: the least weight <= K with a checked witness, and None below it."""
    from tropidist_amd import ordering, sparse_dp
    print("case serve wide-trellis requests", flush=True)
    probs = [(problem.quasi_cyclic(l, m, seed=s), K) for l, m, s, K in WIDE_REQUESTS]
    alone = []
    for p, K in probs:
        order = ordering.greedy_order_native(p.h, p.nrows, p.words())
        with pytest.raises(ValueError):
            [sparse_dp.Trellis.for_query(p, i, order) for i in range(p.nobs)]
        rec = serve.sparse_certify(p, K)
        assert rec["path"] == "search" and rec["witness_ok"]
        if rec["distance"] is not None:
            assert rec["distance"] <= K and len(rec["witness"]) == rec["distance"]
            below = serve.sparse_certify(p, rec["distance"] - 1)
            assert below["distance"] is None
        alone.append(rec)
    srv = serve.Server(memory_bytes=8 << 30, dense_lanes=1, sparse_arena_bytes=1 << 30)
    try:
        tickets = [srv.submit_sparse(p, K) for p, K in probs]
        got = [tk.wait(timeout=300) for tk in tickets]
    finally:
        srv.close()
    for a, g in zip(alone, got):
        assert (g["distance"], g["observable"], g["path"]) == (a["distance"], a["observable"], "search")
        assert g["witness_ok"]
