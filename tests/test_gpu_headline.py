"""The executor the dense benchmark times, pinned at its sizes: fused.run_many (tdp_run_programs, one k_pass_multi
launch per round runs the next pass of every program) on the recorded fixtures whose tables reach 2^26-2^28 entries,
all observables of an instance at once, against the recorded distances; and a weighted four-observable banded model
of the single-large-instance shape at 2^22 entries against the reference's own viterbi.c (compiled from its sources
into oracle/_ref) or the oracle's C."""
import pytest

import oracle
from helpers import dense_fixtures
from tropidist_amd import fused, kernel, problem, trellis

pytestmark = pytest.mark.gpu

WIDE = ("surface_d7", "walkingcat_54", "sdbb_64", "bb72_r6")


@pytest.mark.parametrize("name", WIDE)
def test_run_many_on_wide_fixtures(name):
    for r in dense_fixtures():
        if r["name"] != name:
            continue
        plans = [trellis.build_plan(r["h"], r["l"], 1 << q["observable"], r["order"], r["nrows"]) for q in r["plans"]]
        assert max(pl.peak_bits for pl in plans) >= 26
        print("case run_many", name, r["variant"], "observables", len(plans), "peak",
              max(pl.peak_bits for pl in plans), flush=True)
        got = fused.run_many(plans)
        assert got == list(r["distances"]), (name, got, r["distances"])
        # and the same instance with every other observable as a witness run in the same call
        wit = [i % 2 == 1 for i in range(len(plans))]
        mixed = fused.run_many(plans, wit)
        for i, (g, d) in enumerate(zip(mixed, r["distances"])):
            assert (g[0] if wit[i] else g) == d, (name, i)


def _reference_value(plan):
    return oracle.run_plan_c(plan, "ref" if oracle.have_ref() else "oracle")


def test_weighted_four_observable_banded_at_2_22():
    """Costs 1-3, four logical rows of density 0.05, 2-6 detectors per column in a 22-row window: the programs of all
    four observables in one run_many call equal the CPU reference's value for each observable."""
    p = problem.banded(R=300, N=3000, kmin=2, kmax=6, W=21, nobs=4, rho=0.05, weights=(1, 3), seed=2022)
    order = list(range(p.ncols))
    plans = [trellis.build_plan(p.h, p.l, 1 << obs, order, p.nrows, weights=p.w) for obs in range(p.nobs)]
    plans = [pl for pl in plans if pl.feasible]
    assert len(plans) >= 2 and max(pl.peak_bits for pl in plans) >= 22
    print("case run_many weighted banded 2^%d" % max(pl.peak_bits for pl in plans), flush=True)
    got = fused.run_many(plans)
    want = [_reference_value(pl) for pl in plans]
    assert got == want
    assert [kernel.run_plan(pl, executor="per-op")[0] for pl in plans] == want


def test_witness_of_the_single_large_instance_at_2_26():
    """The benchmark's single large instance (banded R=600 N=6000 W=25, costs 1-3, four logical rows, seed 2000):
    the minimising observable's fused witness run at 2^26 entries gives the step's distance, and its columns form a
    solution of that cost (H e = 0, the observable flipped) -- the certificate the reference's dp_witness gives."""
    from tropidist_amd import auto, workloads
    from tropidist_amd.problem import verify_witness
    p = workloads.single_large(W=25)
    order = list(range(p.ncols))
    plans = trellis.build_plans_native(p.h, p.l, [1 << i for i in range(p.nobs)], order, p.nrows, p.w)
    feas = [(i, pl) for i, pl in enumerate(plans) if pl.feasible]
    values = fused.run_many([pl for _, pl in feas])
    best = min(values)
    obs, plan = feas[values.index(best)]
    assert plan.peak_bits == 26
    print("case witness of the single large instance, observable", obs, "peak 2^26", flush=True)
    v, cols = auto.witness_plan(plan, order)
    assert v == best and verify_witness(p, cols, best, obs) and p.cost(cols) == best
