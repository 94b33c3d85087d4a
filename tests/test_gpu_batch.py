"""GPU batched certification: every problem's certificate from one persistent launch equals the one-at-a-time
result (values per observable) and carries a valid witness of that cost."""
import random

import numpy as np
import pytest

import oracle
from helpers import random_problem
from tropidist_amd import auto, batch, problem, workloads
from tropidist_amd.problem import Problem, verify_witness

pytestmark = pytest.mark.gpu


def test_batch_random_small_problems_against_brute_force():
    print("case batch random small", flush=True)
    rng = random.Random(41)
    probs = []
    for t in range(300):
        m, k, h, l = random_problem(rng)
        w = [rng.randint(1, 3) for _ in h] if t % 2 else None
        probs.append(Problem("r%d" % t, m, k, h, l, w=w))
    orders = [list(range(p.ncols)) for p in probs]
    recs = batch.certify_batch(probs, orders)
    for p, rec in zip(probs, recs):
        per = [oracle.brute(p.h, p.l, 1 << i, p.w) for i in range(p.nobs)]
        assert rec["per_observable"] == per
        if rec["distance"] is not None:
            assert rec["witness_ok"] and rec["upper_bound"] == rec["distance"]


def test_batch_banded_equals_one_at_a_time():
    print("case batch banded vs single", flush=True)
    rng = np.random.default_rng(7)
    probs = [workloads.banded_instance(rng, 7000 + i, int(rng.integers(6, 17)), bool(i % 2)) for i in range(24)]
    recs = batch.certify_batch(probs)
    for p, rec in zip(probs, recs):
        single = auto.certify(p)
        assert rec["per_observable"] == single["per_observable"], p.name
        if rec["distance"] is not None:
            assert rec["witness_ok"] and verify_witness(p, rec["witness"], rec["distance"], rec["observable"])


def test_batch_chain_boundary():
    print("case batch chain", flush=True)
    recs = batch.certify_batch([problem.chain(254), problem.chain(30)], [list(range(254)), list(range(30))])
    assert recs[0]["distance"] == 254 and recs[0]["witness"] == list(range(254))
    assert recs[1]["distance"] == 30
    with pytest.raises(OverflowError):
        batch.certify_batch([problem.chain(255)], [list(range(255))])


@pytest.mark.parametrize("concurrent", [False, True])
def test_hybrid_wide_path_equals_persistent_path(concurrent):
    """Problems split between the persistent launches and the fused executor on worker streams (wide_bits lowered
    so both paths carry work): the same values as the persistent path alone, checked witnesses."""
    print("case batch hybrid concurrent", concurrent, flush=True)
    rng = np.random.default_rng(11)
    probs = [workloads.banded_instance(rng, 7100 + i, int(rng.integers(8, 19)), bool(i % 2)) for i in range(40)]
    orders, flat, owner = batch.plan_problems(probs, None)
    assert min(pl.peak_bits for pl in flat if pl.feasible) < 14 <= max(pl.peak_bits for pl in flat if pl.feasible)
    ref = batch.certify_planned(probs, orders, flat, owner, wide_bits=None)
    got = batch.certify_planned(probs, orders, flat, owner, wide_bits=14, wide_concurrent=concurrent)
    for p, a, b in zip(probs, ref, got):
        assert a["per_observable"] == b["per_observable"], p.name
        if b["distance"] is not None:
            assert b["witness_ok"] and verify_witness(p, b["witness"], b["distance"], b["observable"])
            assert b["upper_bound"] == b["distance"] == b["lower_bound"]


def test_hybrid_overflow_from_a_wide_worker():
    """A cost of 255 or more found on a worker stream raises OverflowError in the caller."""
    p = problem.banded(R=60, N=300, kmin=2, kmax=5, W=10, nobs=1, rho=0.1, seed=5)
    p.w = [200] * p.ncols
    probs = [p, problem.chain(30)]
    orders, flat, owner = batch.plan_problems(probs, [list(range(q.ncols)) for q in probs])
    with pytest.raises(OverflowError):
        batch.certify_planned(probs, orders, flat, owner, wide_bits=4)
