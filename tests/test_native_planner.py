"""The C++ planner (tdp_plan_build) emits the same op streams, widths and work as the Python restatement of the
reference's planner, on random problems (with costs) and on every recorded fixture."""
import random

from helpers import dense_fixtures, random_problem
from tropidist_amd import trellis, workloads


def _same(a, b):
    assert a.feasible == b.feasible
    if a.feasible:
        assert a.ops == b.ops and a.peak_bits == b.peak_bits and a.work() == b.work() and a.columns == b.columns
        assert (a.encode() == b.encode()).all()


def test_random_problems():
    rng = random.Random(1)
    for t in range(2000):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        w = [rng.randint(1, 3) for _ in h] if t % 2 else None
        masks = [rng.randrange(1, 1 << k) for _ in range(2)]
        native = trellis.build_plans_native(h, l, masks, order, m, weights=w)
        for om, b in zip(masks, native):
            _same(trellis.build_plan(h, l, om, order, m, weights=w), b)


def test_recorded_fixtures_and_banded():
    for r in dense_fixtures():
        masks = [1 << p["observable"] for p in r["plans"]]
        for om, b in zip(masks, trellis.build_plans_native(r["h"], r["l"], masks, r["order"], r["nrows"])):
            _same(trellis.build_plan(r["h"], r["l"], om, r["order"], r["nrows"]), b)
    p = workloads.single_cpu()
    order = list(range(p.ncols))
    _same(trellis.build_plan(p.h, p.l, 1, order, p.nrows), trellis.build_plans_native(p.h, p.l, [1], order, p.nrows)[0])


def test_reference_style_plan_objects_encode_identically():
    from tropidist_amd import problem
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=3)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)

    class Ref:
        ops = [tuple(op) for op in plan.ops]
        peak_bits = plan.peak_bits
        feasible = True
    q = trellis.as_plan(Ref())
    assert (q.encode() == plan.encode()).all() and q.work() == plan.work() and q.peak_bits == plan.peak_bits
    assert trellis.as_plan(plan) is plan


def test_native_cut_profile_equals_python():
    """tdp_rank_profile gives the cut ranks of trellis.cut_profile (random problems, multi-word rows, banded)."""
    from tropidist_amd import problem
    rng = random.Random(9)
    for t in range(200):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        assert trellis.cut_profile_native(h, order, m) == trellis.cut_profile(h, order), t
    for R, W in ((150, 12), (300, 20)):
        p = problem.banded(R=R, N=4 * R, kmin=2, kmax=6, W=W, nobs=2, rho=0.1, seed=R)
        order = list(range(p.ncols))
        assert trellis.cut_profile_native(p.h, order, p.nrows) == trellis.cut_profile(p.h, order)
        rev = order[::-1]
        assert trellis.cut_profile_native(p.h, rev, p.nrows) == trellis.cut_profile(p.h, rev)


def _sparse_cases():
    from sparse_helpers import random_problem as sparse_random_problem
    from tropidist_amd import problem
    rng = random.Random(11)
    out = [sparse_random_problem(rng) for _ in range(300)]
    out += [problem.quasi_cyclic(l, m, seed=s) for l, m, s in
            [(6, 6, 3001), (6, 8, 3003), (9, 8, 3004), (12, 12, 3000), (15, 9, 5012), (6, 15, 5010)]]
    return rng, out


def test_greedy_order_native_equals_python():
    """tdp_greedy_order gives greedy_straddle_order's column order exactly (random problems, QC codes, any start)."""
    from tropidist_amd import ordering
    rng, probs = _sparse_cases()
    for p in probs:
        start = rng.randrange(p.ncols) if p.ncols else 0
        assert ordering.greedy_order_native(p.h, p.nrows, start=start) == \
            ordering.greedy_straddle_order(p.h, p.nrows, start=start)


def _target_row(p, q, order):
    return sum(((p.l[j] >> q) & 1) << pos for pos, j in enumerate(order))


def test_sparse_trellis_native_equals_oracle_restatement():
    """tdp_sp_trellis (sparse_dp.Trellis and the engine's query_words) gives, per observable, the feasibility, slot
    count, engine words, active rows and widths of the oracle's restatement of the reference's trellis (random
    problems in random orders and QC codes in the greedy order), and the same slot-limit refusal."""
    import numpy as np

    import oracle
    from tropidist_amd import ordering, sparse_batch, sparse_dp
    rng, probs = _sparse_cases()
    checked = 0
    for p in probs:
        order = list(range(p.ncols))
        if p.ncols > 100:
            order = ordering.greedy_straddle_order(p.h, p.nrows)
        else:
            rng.shuffle(order)
        want, too_wide = [], None
        for q in range(p.nobs):
            try:
                ref = oracle.SparseTrellis(p.h, _target_row(p, q, order), order, p.nrows)
            except ValueError as exc:
                too_wide = str(exc)
                break
            if ref.feasible:
                want.append((q, ref) + sparse_batch.trellis_words(ref))
        if too_wide is not None:
            for call in (lambda: sparse_batch.query_words(p, order),
                         lambda: [sparse_dp.Trellis.for_query(p, q, order) for q in range(p.nobs)]):
                try:
                    call()
                except ValueError as exc:
                    assert str(exc) == too_wide
                    continue
                raise AssertionError("native planner accepted a trellis past the slot limit")
            continue
        got = sparse_batch.query_words(p, order)
        assert [g[0] for g in got] == [w[0] for w in want]
        for (q, prm, tm), (_, ref, wprm, wtm) in zip(got, want):
            assert np.array_equal(prm[:p.ncols], wprm[:p.ncols]) and np.array_equal(tm, wtm), q
            tr = sparse_dp.Trellis.for_query(p, q, order)
            assert tr.nslots == ref.nslots and tr.width == ref.width and tr.tmask == ref.tmask
            assert tr.active == [sum(1 << ref.slot[r] for r in rows) for rows in ref.active_rows]
            checked += 1
    assert checked > 200


def test_lookahead_bounds_native_equal_oracle_restatement():
    """tdp_sp_bounds gives the oracle restatement's lookahead classes and extras at every cut, both directions."""
    import oracle
    from tropidist_amd import ordering, sparse_dp
    rng, probs = _sparse_cases()
    checked = 0
    for p in probs[:120] + probs[-6:]:
        order = list(range(p.ncols))
        if p.ncols > 100:
            order = ordering.greedy_straddle_order(p.h, p.nrows)
        else:
            rng.shuffle(order)
        for q in range(p.nobs):
            try:
                ref = oracle.SparseTrellis(p.h, _target_row(p, q, order), order, p.nrows)
            except ValueError:
                continue
            if not ref.feasible:
                continue
            tr = sparse_dp.Trellis.for_query(p, q, order)
            for fwd in (True, False):
                assert tr.bound_data(fwd) == ref.bound_data(fwd), (q, fwd)
            checked += 1
    assert checked > 100


def test_planning_order_is_best_orders_choice():
    from tropidist_amd import auto, problem
    for seed in range(4):
        p = problem.banded(R=60, N=240, kmin=2, kmax=4, W=8 + seed, nobs=2, rho=0.1, seed=seed)
        assert auto.planning_order(p) == auto.best_order(p, seeds=0)[0]


def test_more_than_63_active_rows_is_refused():
    # a dense random 80 x 200 matrix: about 80 rows are active in the middle of any order
    import pytest
    rng = random.Random(5)
    nrows, ncols = 80, 200
    h = [rng.getrandbits(nrows) | 1 for _ in range(ncols)]
    l = [rng.getrandbits(1) for _ in range(ncols)]
    with pytest.raises(ValueError):
        trellis.build_plans_native(h, l, [1], list(range(ncols)), nrows)
