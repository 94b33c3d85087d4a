"""Pins the dense oracle (oracle/oracle.py, oracle/tdp_oracle.c) and the planner before they check the GPU:
brute force (tests/test_dp.py pattern, with the reference's recorded count of feasible problems), the reference's
own viterbi.c compiled into oracle/_ref, the chain closed form at the 8-bit boundary and the golden fixtures."""
import random

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures, random_problem
from tropidist_amd import problem, trellis

needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built (make -C oracle)")


def test_random_problems_against_brute_force():
    """3000 problems, seed 1: every executor equals brute force; 1966 feasible and widest table 2^7, as the
    reference recorded (results/dense_audit_20261021/kernel_validation.txt:1)."""
    rng = random.Random(1)
    feasible = max_bits = 0
    impls = ["oracle"] + (["ref"] if oracle.have_ref() else [])
    for t in range(3000):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        obs_mask = rng.randrange(1, 1 << k)
        want = oracle.brute(h, l, obs_mask)
        plan = trellis.build_plan(h, l, obs_mask, order, m)
        assert oracle.run_plan_numpy(plan) == want, t
        for impl in impls:
            assert oracle.run_plan_c(plan, impl) == want, (t, impl)
        if plan.feasible:
            cols = [h[j] | ((bin(l[j] & obs_mask).count("1") & 1) << m) for j in range(len(h))]
            lam = trellis.cut_profile(cols, order)
            lam_h = trellis.cut_profile(h, order)
            for i, w in enumerate(plan.widths):
                assert lam_h[i + 1] <= w <= lam_h[i + 1] + 1 and w >= lam[i + 1]
            max_bits = max(max_bits, plan.peak_bits)
        feasible += want is not None
    assert feasible == 1966 and max_bits == 7


def test_big_tables_numpy_equals_c():
    """tests/test_dp.py:65-78 of the reference (seed 7, 40 problems, tables >= 128 entries)."""
    rng = random.Random(7)
    for t in range(40):
        m = rng.randint(10, 16)
        n = rng.randint(m + 4, m + 14)
        h = [rng.randrange(1, 1 << m) for _ in range(n)]
        l = [rng.randrange(0, 2) for _ in range(n)]
        plan = trellis.build_plan(h, l, 1, list(range(n)), m)
        a = oracle.run_plan_numpy(plan)
        assert oracle.run_plan_c(plan, "oracle") == a
        if oracle.have_ref():
            assert oracle.run_plan_c(plan, "ref") == a


def test_weighted_problems_against_brute_force():
    rng = random.Random(11)
    for t in range(600):
        m, k, h, l = random_problem(rng)
        w = [rng.randint(1, 3) for _ in h]
        order = list(range(len(h)))
        rng.shuffle(order)
        obs_mask = rng.randrange(1, 1 << k)
        want = oracle.brute(h, l, obs_mask, w)
        plan = trellis.build_plan(h, l, obs_mask, order, m, weights=w)
        assert oracle.run_plan_numpy(plan) == want, t
        assert oracle.run_plan_c(plan, "oracle") == want, t
        if oracle.have_ref():
            assert oracle.run_plan_c(plan, "ref") == want, t


@needs_ref
@pytest.mark.parametrize("bits", [0, 1, 3, 4, 5, 6, 7, 9, 12])
def test_per_op_kernels_equal_reference_c(bits):
    """Each oracle kernel equals the reference's compiled viterbi.c on random tables (all masks for small tables,
    random masks with low and high top bits for larger ones), including saturation at 255."""
    rng = np.random.default_rng(bits)
    O, R = oracle.olib(), oracle.rlib("viterbi")
    n = 1 << bits
    masks = list(range(1, n)) if n <= 64 else sorted({int(x) for x in rng.integers(1, n, size=40)} | {1, n - 1, n >> 1})
    for h in masks:
        for w in (1, 2, 3, 200):
            T = rng.integers(0, 256, size=n, dtype=np.uint8)
            T[rng.random(n) < 0.3] = 255
            a, b = T.copy(), T.copy()
            if w == 1:
                O.or_dp_update(oracle.ptr(a), n, h)
                R.dp_update(oracle.ptr(b), n, h)
            else:
                O.or_dp_update_w(oracle.ptr(a), n, h, w)
                R.dp_update_w(oracle.ptr(b), n, h, w)
            assert np.array_equal(a, b), (n, h, w)
    for hl in masks[:20] + [0]:
        if hl >= n:
            continue
        T = rng.integers(0, 256, size=2 * n, dtype=np.uint8)
        a, b = T.copy(), T.copy()
        O.or_dp_expand_update(oracle.ptr(a), n, hl)
        R.dp_expand_update(oracle.ptr(b), n, hl)
        assert np.array_equal(a, b), (n, hl)
    for pos in range(max(1, bits)):
        for target in (0, 1):
            T = rng.integers(0, 256, size=n, dtype=np.uint8)
            a, b = T.copy(), T.copy()
            assert O.or_dp_select(oracle.ptr(a), n, pos, target) == R.dp_select(oracle.ptr(b), n, pos, target)
            assert np.array_equal(a[:max(1, n // 2)], b[:max(1, n // 2)]), (n, pos, target)
    T = rng.integers(0, 256, size=n, dtype=np.uint8)
    ha, hb = np.zeros(256, np.uint64), np.zeros(256, np.uint64)
    O.or_dp_hist(oracle.ptr(T), n, oracle.ptr(ha))
    R.dp_hist(oracle.ptr(T), n, oracle.ptr(hb))
    assert np.array_equal(ha, hb) and int(ha.sum()) == n


def test_chain_at_the_8bit_boundary():
    """tests/test_false_bounds.py:184-203: chain(n) has distance n; 254 is exact, 255 and 256 raise OverflowError;
    a query without a logical is None.  Weighted chains cost the sum of their weights."""
    for n in (254, 255, 256):
        p = problem.chain(n)
        plan = trellis.build_plan(p.h, p.l, 1, list(range(n)), p.nrows)
        assert plan.feasible
        for run in (oracle.run_plan_numpy, lambda q: oracle.run_plan_c(q, "oracle")):
            if n < 255:
                assert run(plan) == n
            else:
                with pytest.raises(OverflowError):
                    run(plan)
    p = problem.Problem("no logical", 1, 1, [1, 1], [1, 1])
    plan = trellis.build_plan(p.h, p.l, 1, [0, 1], p.nrows)
    assert not plan.feasible and oracle.run_plan_numpy(plan) is None
    rng = random.Random(4)
    for n in (5, 40, 90):
        w = [rng.randint(1, 3) for _ in range(n)]
        p = problem.chain(n, weights=w)
        plan = trellis.build_plan(p.h, p.l, 1, list(range(n)), p.nrows, weights=w)
        want = sum(w)
        if want < 255:
            assert oracle.run_plan_numpy(plan) == want == oracle.run_plan_c(plan, "oracle")
        else:
            with pytest.raises(OverflowError):
                oracle.run_plan_c(plan, "oracle")


def test_planner_matches_the_recorded_plans():
    """peak_bits and work of every observable's plan, for the recorded column order, equal the reference's
    (results/dense_audit_20261021/widths)."""
    for r in dense_fixtures():
        for p in r["plans"]:
            plan = trellis.build_plan(r["h"], r["l"], 1 << p["observable"], r["order"], r["nrows"])
            assert plan.feasible == p["feasible"], (r["name"], p)
            if plan.feasible:
                assert (plan.peak_bits, plan.work()) == (p["peak_bits"], p["work"]), (r["name"], p)


def test_oracle_reproduces_recorded_distances():
    """Per-observable distances of the recorded runs (results/dense_audit_20261021/runs) for the fixtures whose
    tables are at most 2^18 entries."""
    checked = 0
    for r in dense_fixtures():
        if r["distances"] is None or r["peak_bits"] > 18:
            continue
        for p, d in zip(r["plans"], r["distances"]):
            plan = trellis.build_plan(r["h"], r["l"], 1 << p["observable"], r["order"], r["nrows"])
            assert oracle.run_plan_c(plan, "oracle") == d, (r["name"], p["observable"])
            checked += 1
    assert checked >= 20
