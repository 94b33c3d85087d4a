"""GPU witnesses of the dense DP: every returned column set is a solution (H e = 0, the observable flipped) whose
cost equals the brute-force minimum (tests/test_auto.py:79-87 and test_false_bounds.py:184-203 of the reference)."""
import random

import pytest

import oracle
from helpers import dense_fixtures, random_problem
from tropidist_amd import auto, problem, trellis
from tropidist_amd.problem import Problem, verify_witness

pytestmark = pytest.mark.gpu


def test_random_witnesses_against_brute_force():
    print("case witness random seed 21", flush=True)
    rng = random.Random(21)
    for t in range(500):
        m, k, h, l = random_problem(rng)
        w = [rng.randint(1, 3) for _ in h] if t % 2 else None
        order = list(range(len(h)))
        rng.shuffle(order)
        obs = rng.randrange(k)
        p = Problem("r", m, k, h, l, w=w)
        want = oracle.brute(h, l, 1 << obs, w)
        cols = auto.dp_witness(p, order, obs)
        if want is None:
            assert cols is None
        else:
            assert verify_witness(p, cols, want, obs), (t, cols, want)


def test_chain_witness_is_the_whole_chain():
    print("case witness chain", flush=True)
    for n in (20, 254):
        p = problem.chain(n)
        assert auto.dp_witness(p, list(range(n)), 0) == list(range(n))
    with pytest.raises(OverflowError):
        auto.dp_witness(problem.chain(255), list(range(255)), 0)


def test_certificates_of_recorded_instances():
    for r in dense_fixtures():
        if r["distances"] is None or r["peak_bits"] > 24:
            continue
        print("case certify", r["name"], r["variant"], flush=True)
        p = Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"])
        rec = auto.certify(p, r["order"])
        want = min(d for d in r["distances"] if d is not None)
        assert rec["distance"] == rec["lower_bound"] == rec["upper_bound"] == want and rec["witness_ok"]


def test_weighted_banded_certificate():
    print("case certify banded W18 weighted", flush=True)
    p = problem.banded(R=100, N=500, kmin=2, kmax=6, W=18, nobs=3, rho=0.05, weights=(1, 3), seed=4242)
    order = list(range(p.ncols))
    rec = auto.certify(p, order)
    for obs in range(p.nobs):
        plan = trellis.build_plan(p.h, p.l, 1 << obs, order, p.nrows, weights=p.w)
        want = oracle.run_plan_c(plan, "oracle") if plan.feasible else None
        assert rec["per_observable"][obs] == want
    assert rec["witness_ok"] and rec["upper_bound"] == rec["distance"]
