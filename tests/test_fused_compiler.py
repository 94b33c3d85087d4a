"""The fused-pass compiler on the CPU: simulate() executes every pass with the kernel's exact index arithmetic
(input / output addresses, new-row fills, selected-row store masks, in-dword lane permutations) and must give the
oracle's value on every plan, and the recorded distances of the reference's audit runs."""
import random

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures, random_problem
from tropidist_amd import fused, problem, trellis


def _random_plan(rng, t):
    m = rng.randint(4, 12)
    n = rng.randint(6, 30)
    h = [rng.randrange(1, 1 << m) for _ in range(n)]
    l = [rng.randrange(0, 2) for _ in range(n)]
    w = [rng.randint(1, 3) for _ in range(n)] if t % 2 else None
    order = list(range(n))
    rng.shuffle(order)
    return trellis.build_plan(h, l, 1, order, m, weights=w)


@pytest.mark.parametrize("dmax", [1, 3, 5])
def test_random_plans_match_oracle(dmax):
    rng = random.Random(10 + dmax)
    checked = passes = 0
    for t in range(150):
        plan = _random_plan(rng, t)
        if not plan.feasible:
            continue
        prog = fused.compile_program(plan, dmax=dmax, min_bits=2)
        passes += len(prog.passes)
        assert all(p.D <= dmax for p in prog.passes)
        assert fused.simulate(prog, plan) == oracle.run_plan_numpy(plan), t
        checked += 1
    assert checked > 80 and passes > 200


def test_reference_random_problems_against_brute_force():
    """The generator of the reference's tests/test_dp.py:26-33 (seed 1, first 400): fused value == brute force."""
    rng = random.Random(1)
    for t in range(400):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        obs_mask = rng.randrange(1, 1 << k)
        plan = trellis.build_plan(h, l, obs_mask, order, m)
        if not plan.feasible:
            continue
        got = fused.simulate(fused.compile_program(plan, min_bits=2), plan)
        assert got == oracle.brute(h, l, obs_mask), t


def test_recorded_distances():
    """Per-observable distances of the reference's recorded audit runs (tables up to 2^16 entries)."""
    checked = 0
    for r in dense_fixtures():
        if r["distances"] is None or r["peak_bits"] > 16:
            continue
        for p, d in zip(r["plans"], r["distances"]):
            plan = trellis.build_plan(r["h"], r["l"], 1 << p["observable"], r["order"], r["nrows"])
            prog = fused.compile_program(plan, min_bits=6)
            assert fused.simulate(prog, plan) == d, (r["name"], p["observable"])
            checked += 1
    assert checked >= 10


def test_banded_weighted_layouts():
    """Wider banded plans: both store modes occur, every pass has at most DMAX directions, and the value equals
    the oracle's C executor."""
    for W, seed in ((8, 1), (12, 2)):
        p = problem.banded(R=60, N=400, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=(1, 3), seed=seed)
        plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=p.w)
        prog = fused.compile_program(plan, min_bits=4)
        ps = prog.passes
        assert ps and all(q.D <= fused.DMAX for q in ps)
        assert any(q.byte_store for q in ps) and any(not q.byte_store for q in ps)
        # every op of the stream is covered once, in order
        covered = []
        for s in prog.segments:
            covered.append((s[1], s[2]) if s[0] == "ops" else (s[1].first, s[1].last))
        assert covered[0][0] == 0 and covered[-1][1] == len(plan.encode())
        assert all(a[1] == b[0] for a, b in zip(covered, covered[1:]))
        assert fused.simulate(prog, plan) == oracle.run_plan_c(plan, "oracle")


def test_pack_layout():
    """Descriptor records: size of the C struct (880 bytes), fields where the executor reads them."""
    assert fused.PASS_DTYPE.itemsize == 880
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=5)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    prog = fused.compile_program(plan, min_bits=4)
    segs, arr = fused.pack(prog)
    assert segs.shape == (len(prog.segments), 3)
    for s, seg in zip(segs, prog.segments):
        if seg[0] == "ops":
            assert s[0] == fused.SEG_OPS
        else:
            q = seg[1]
            r = arr[s[1]]
            assert s[0] == fused.SEG_PASS
            assert (r["D"], r["nfree"], r["b_in"], r["b_out"]) == (q.D, q.nfree, q.b_in, q.b_out)
            assert list(r["din"][:q.D + 2]) == q.din
            for m, runs in enumerate(q.runs()):
                assert r["nruns"][m] == len(runs)
                got = [fused.unpack_run(r["runs"][m, i]) for i in range(len(runs))]
                assert got == runs
            # the runs reproduce the per-variable maps
            tid = np.arange(1 << q.nfree, dtype=np.int64)
            want = fused._xor_fold(q.fout, tid)
            assert (fused._map_index(q.runs()[1], tid) == want).all()
            assert r["byte_store"] == int(q.byte_store)


def test_library_descriptor_size():
    from tropidist_amd import _lib
    assert int(_lib.lib().tdp_pass_args_bytes()) == fused.PASS_DTYPE.itemsize


def _cols(plan, order, used):
    ordinal = np.cumsum(plan.encode()[:, 0] == 1) - 1
    return sorted(order[plan.columns[int(ordinal[i])]] for i in used)


def test_witness_mode_traceback_on_cpu():
    """The witness-mode program and the traceback arithmetic of the device kernel give a minimum-cost solution:
    value == brute force, and the columns satisfy H e = 0, (L e)_obs = 1 with that cost (random problems of the
    reference's generator, half with costs 1-3, and weighted banded models)."""
    from tropidist_amd.problem import Problem, verify_witness
    rng = random.Random(5)
    checked = 0
    for t in range(250):
        m, k, h, l = random_problem(rng)
        w = [rng.randint(1, 3) for _ in h] if t % 2 else None
        order = list(range(len(h)))
        rng.shuffle(order)
        plan = trellis.build_plan(h, l, 1, order, m, weights=w)
        if not plan.feasible:
            continue
        v, used = fused.witness_cpu(plan)
        assert v == oracle.brute(h, l, 1, w), t
        assert verify_witness(Problem("r", m, k, h, l, w=w), _cols(plan, order, used), v, 0), t
        checked += 1
    assert checked > 120
    for W, seed in ((8, 1), (12, 2)):
        p = problem.banded(R=60, N=400, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=(1, 3), seed=seed)
        order = list(range(p.ncols))
        plan = trellis.build_plan(p.h, p.l, 1, order, p.nrows, weights=p.w)
        prog, steps, _ = fused.witness_program(plan)
        assert all(len(q.upd_ops) <= fused.WITNESS_MAX_UPDATES for q in prog.passes)
        v, used = fused.witness_cpu(plan)
        assert v == oracle.run_plan_c(plan, "oracle")
        assert verify_witness(p, _cols(plan, order, used), v, 0)


def test_witness_step_layout():
    from tropidist_amd import _lib
    assert fused.WSTEP_DTYPE.itemsize == int(_lib.lib().tdp_witness_step_bytes()) == 240
