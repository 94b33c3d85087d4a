"""The front end as the reference's own callers use it, on the GPU:

* the sparse entry points as bench/adapters/ours_pipeline.py:675-714 calls them (sparse_trellises, sparse_estimate
  with trellises / max_bytes / max_seconds, sparse_distance with trellises / schedule / lower / deadline / max_bytes),
  against the dense DP's distance on the same codes;
* the device tables behind the reference's backend interface: HipTables.step with the lookahead data in the
  reference's host form (Trellis.bound_data triples, what cdist/sparse_dp.py:649 passes to a non-"c" backend), and
  query() through a backend object (one column step per call) equal to the native runner, per-cut counts included;
* kernel.run_plan_numpy (the reference's name) equal to run_plan."""
import random
import time

import pytest

from sparse_helpers import brute, check_witness, random_problem
from tropidist_amd import auto, kernel, problem, sparse_dp, trellis

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [3001, 3003])
def test_pipeline_sparse_calls_equal_dense_distance(seed):
    p = problem.quasi_cyclic(6, 6, seed=seed)
    order = auto.best_order(p, seeds=0)[0]
    d, per = auto.dp_distance(p, order, max_table_bytes=1 << 34, deadline=time.monotonic() + 600)
    assert d is not None
    print("case pipeline sparse seed", seed, "dense distance", d, flush=True)
    trs = auto.sparse_trellises(p, order)
    budget = 8 << 30
    for K, lower in ((d - 1, 1), (d, d - 1), (d + 1, 1)):
        est = auto.sparse_estimate(p, K, order, trellises=trs, max_bytes=budget, max_seconds=120.0)
        assert est["fits"] and est["queries"] == sum(t.feasible for t in trs)
        assert len(est["schedule"]) == p.ncols + 1 and est["predicted_seconds"] > 0
        best = auto.sparse_distance(p, K, order, trellises=trs, schedule=est["schedule"], lower=lower,
                                    deadline=time.monotonic() + 600, max_bytes=budget)
        if K < d:
            assert best is None
        else:
            value, cols, obs = best
            assert value == d and per[obs] == d
            check_witness(p, cols, d, obs)


def test_sparse_distance_memory_and_time_limits():
    p = problem.quasi_cyclic(6, 6, seed=3001)
    order = auto.best_order(p, seeds=0)[0]
    with pytest.raises(MemoryError):
        auto.sparse_distance(p, 8, order, max_bytes=48 * 8)
    with pytest.raises(TimeoutError):
        auto.sparse_distance(p, 8, order, deadline=time.monotonic() - 1.0)


def test_hip_step_takes_reference_bound_triples():
    print("case hip step with bound triples", flush=True)
    rng = random.Random(12)
    checked = 0
    while checked < 25:
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        tr = sparse_dp.Trellis.for_query(p, 0, order)
        if not tr.feasible:
            continue
        be = sparse_dp.HipTables(tr)
        triples, dev = tr.bound_data(True), be.bound(True)
        tab = be.start()
        K = p.ncols
        for pos in range(tr.N):
            c = pos + 1
            a, _ = be.step(tab, tr.col[pos], tr.fsel_mask[pos], tr.fsel_bits[pos], K, K, tr.tmask[c], triples[c])
            b, _ = be.step(tab, tr.col[pos], tr.fsel_mask[pos], tr.fsel_bits[pos], K, K, tr.tmask[c], dev[c])
            ka, va = be.arrays(a)
            kb, vb = be.arrays(b)
            assert sorted(zip(ka.tolist(), va.tolist())) == sorted(zip(kb.tolist(), vb.tolist()))
            tab = a
        checked += 1


def test_query_through_backend_object_equals_native_runner():
    """The module's Sweep stepping HipTables column by column and the native runner give the same value, the same
    per-cut counts and valid witnesses, with and without the lookahead, forward and meet in the middle."""
    print("case backend object vs native", flush=True)
    rng = random.Random(21)
    checked = 0
    for _ in range(40):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        tr = sparse_dp.Trellis.for_query(p, 0, order)
        if not tr.feasible:
            continue
        truth = brute(p, 0)
        for K in range(1, p.ncols + 1):
            want = truth if truth <= K else None
            for mode, split in (("forward", None), ("mitm", None), ("mitm", "schedule")):
                use_h = rng.random() < 0.5
                a = sparse_dp.query(tr, K, mode=mode, split=split, use_h=use_h, backend="hip")
                b = sparse_dp.query(tr, K, mode=mode, split=split, use_h=use_h, backend=sparse_dp.HipTables(tr))
                assert a["value"] == b["value"] == want, (mode, K, a["value"], b["value"], want)
                for key in ("forward_counts", "backward_counts", "changed_counts"):
                    if key in a:
                        assert a[key] == b[key], key
                if want is not None:
                    check_witness(p, a["witness"], want, 0)
                    check_witness(p, b["witness"], want, 0)
                checked += 1
    assert checked > 100


def test_run_plan_numpy_name_runs_on_the_device():
    p = problem.banded(R=60, N=300, kmin=1, kmax=4, W=10, nobs=1, rho=0.1, seed=1000)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    assert kernel.run_plan_numpy(plan) == kernel.run_plan(plan)[0]
    infeasible = trellis.build_plan([1, 1], [0, 0], 1, [0, 1], 1)
    assert kernel.run_plan_numpy(infeasible) is None
