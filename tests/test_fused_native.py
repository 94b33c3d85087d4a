"""The C++ fused compiler (tdp_fused_compile) emits exactly the arrays of the Python compiler (the readable statement
of the rules): segments, pass descriptors byte for byte, and witness steps byte for byte."""
import random

import numpy as np

from helpers import dense_fixtures, random_problem
from tropidist_amd import fused, problem, trellis


def _same(plan, **kw):
    segs, arr = fused.pack(fused.compile_program(plan, **kw))
    nsegs, narr = fused.compile_native(plan, **kw)
    assert np.array_equal(segs, nsegs)
    npass = int((segs[:, 0] == fused.SEG_PASS).sum())
    assert arr[:npass].tobytes() == narr[:npass].tobytes()


def _same_witness(plan):
    prog, steps, _ = fused.witness_program(plan)
    segs, arr = fused.pack(prog)
    nsegs, narr, nsteps = fused.compile_native(plan, min_bits=2, max_updates=fused.WITNESS_MAX_UPDATES, witness=True)
    assert np.array_equal(segs, nsegs)
    npass = len(prog.passes)
    assert arr[:npass].tobytes() == narr[:npass].tobytes()
    assert steps.tobytes() == nsteps.tobytes()


def test_random_plans():
    rng = random.Random(21)
    n = 0
    for t in range(300):
        m = rng.randint(4, 12)
        k = rng.randint(6, 30)
        h = [rng.randrange(1, 1 << m) for _ in range(k)]
        l = [rng.randrange(0, 2) for _ in range(k)]
        w = [rng.randint(1, 300) for _ in range(k)] if t % 2 else None
        order = list(range(k))
        rng.shuffle(order)
        plan = trellis.build_plan(h, l, 1, order, m, weights=w)
        if not plan.feasible:
            continue
        for dmax, mb in ((1, 2), (3, 3), (5, 2), (5, 4)):
            _same(plan, dmax=dmax, min_bits=mb)
        _same_witness(plan)
        n += 1
    assert n > 150


def test_reference_generator_plans():
    rng = random.Random(1)
    for t in range(300):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        plan = trellis.build_plan(h, l, rng.randrange(1, 1 << k), order, m)
        if plan.feasible:
            _same(plan, min_bits=2)
            _same_witness(plan)


def test_banded_and_fixtures():
    for W, seed in ((8, 1), (14, 2), (20, 3), (25, 4)):
        p = problem.banded(R=120, N=600, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=(1, 3), seed=seed)
        plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=p.w)
        _same(plan)
        _same(plan, min_bits=4, max_updates=3)
        _same_witness(plan)
    for r in dense_fixtures():
        if r["peak_bits"] > 30:
            continue
        plan = trellis.build_plan(r["h"], r["l"], 1 << r["plans"][0]["observable"], r["order"], r["nrows"])
        if plan.feasible:
            _same(plan)
            _same_witness(plan)
