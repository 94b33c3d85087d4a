"""The front end's dense estimate as the reference's pipeline calls it (bench/adapters/ours_pipeline.py:421:
estimate(p, threads=..., seeds=..., time_limit=..., max_table_bytes=...)): nothing is run, the one-bit margin for
several observables, the early fits=False exit before any plan, and the greedy candidates of best_order
(cdist/auto.py:43-102; the reference's tests/test_auto.py:54-76 check the same rules)."""
import math
from unittest.mock import patch

from tropidist_amd import auto, problem, trellis


def test_estimate_takes_the_pipelines_keywords_and_adds_a_bit_for_several_observables():
    p = problem.banded(R=60, N=240, kmin=2, kmax=4, W=9, nobs=3, rho=0.2, seed=7)
    est = auto.estimate(p, threads=8, seeds=2, time_limit=5.0, max_table_bytes=1 << 30)
    plans = [pl for _, pl in auto._plans(p, est["order"]) if pl.feasible]
    assert est["fits"] and est["queries"] == 3
    assert est["peak_bits"] == plans[0].peak_bits + 1 and est["table_bytes"] == 1 << est["peak_bits"]
    assert est["work"] == plans[0].work() * 3
    assert math.isclose(est["predicted_seconds"], auto.predicted_seconds(est["work"]))
    one = problem.banded(R=60, N=240, kmin=2, kmax=4, W=9, nobs=1, rho=0.2, seed=7)
    e1 = auto.estimate(one, threads=1, seeds=0)
    assert e1["peak_bits"] == [pl for _, pl in auto._plans(one, e1["order"])][0].peak_bits


def test_estimate_refuses_by_cut_rank_before_planning():
    p = problem.banded(R=80, N=400, kmin=2, kmax=5, W=16, nobs=2, rho=0.1, seed=3)
    width = max(trellis.cut_profile(p.h, list(range(p.ncols))))
    with patch.object(auto, "_plans", side_effect=AssertionError("planned")):
        est = auto.estimate(p, threads=1, seeds=0, max_table_bytes=(1 << width) // 4)
    assert not est["fits"] and est["work"] is None and est["predicted_seconds"] == math.inf
    assert est["peak_bits"] == est["cut_rank"] == width


def test_estimate_fits_flag_follows_the_table_budget():
    p = problem.banded(R=60, N=240, kmin=2, kmax=4, W=9, nobs=2, rho=0.2, seed=9)
    est = auto.estimate(p, seeds=0)
    assert auto.estimate(p, seeds=0, max_table_bytes=est["table_bytes"])["fits"]
    if est["peak_bits"] > est["cut_rank"]:
        assert not auto.estimate(p, seeds=0, max_table_bytes=est["table_bytes"] - 1)["fits"]


def test_best_order_candidates_never_lose_to_the_natural_order():
    p = problem.quasi_cyclic(6, 6, seed=3001)
    nat, cost_nat = auto.best_order(p, seeds=0)
    best, cost = auto.best_order(p, seeds=3, time_limit=10.0)
    assert nat == list(range(p.ncols)) and cost <= cost_nat
    assert sorted(best) == list(range(p.ncols))
    assert cost[0] == max(trellis.cut_profile(p.h, best))
