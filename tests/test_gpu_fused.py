"""GPU parity of the fused-pass executor (tdp_run_program): the value of every plan equals the oracle's, bit for bit,
for every pass shape the compiler makes (1-5 directions, in-dword updates, dword and byte stores, passes from 2^2
entries up), and equals the per-op executor at the full 2^25-entry workload."""
import ctypes
import random

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures, random_problem
from tropidist_amd import _lib, fused, kernel, problem, trellis

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dmax,min_bits", [(1, 2), (2, 3), (3, 2), (5, 2), (5, 4)])
def test_random_plans(dmax, min_bits):
    print("case fused random dmax", dmax, "min_bits", min_bits, flush=True)
    rng = random.Random(40 + dmax + 7 * min_bits)
    checked = 0
    for t in range(120):
        m = rng.randint(4, 12)
        n = rng.randint(6, 30)
        h = [rng.randrange(1, 1 << m) for _ in range(n)]
        l = [rng.randrange(0, 2) for _ in range(n)]
        w = [rng.randint(1, 3) for _ in range(n)] if t % 2 else None
        order = list(range(n))
        rng.shuffle(order)
        plan = trellis.build_plan(h, l, 1, order, m, weights=w)
        if not plan.feasible:
            continue
        prog = fused.compile_program(plan, dmax=dmax, min_bits=min_bits)
        assert fused.run_program(plan, prog) == oracle.run_plan_numpy(plan), t
        checked += 1
    assert checked > 60


def test_reference_random_problems_against_brute_force():
    """The generator of the reference's tests/test_dp.py:26-33, seed 1: fused GPU value == brute force."""
    rng = random.Random(1)
    checked = 0
    for t in range(300):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        obs_mask = rng.randrange(1, 1 << k)
        plan = trellis.build_plan(h, l, obs_mask, order, m)
        if not plan.feasible:
            continue
        prog = fused.compile_program(plan, min_bits=2)
        assert fused.run_program(plan, prog) == oracle.brute(h, l, obs_mask), t
        checked += 1
    assert checked > 100


def test_recorded_distances():
    """Distances of the reference's recorded audit runs, tables up to 2^22 entries, default compilation."""
    checked = 0
    for r in dense_fixtures():
        if r["distances"] is None or r["peak_bits"] > 22:
            continue
        for p, d in zip(r["plans"], r["distances"]):
            print("case fused fixture", r["name"], p["observable"], flush=True)
            plan = trellis.build_plan(r["h"], r["l"], 1 << p["observable"], r["order"], r["nrows"])
            assert fused.run_program(plan) == d, (r["name"], p["observable"])
            checked += 1
    assert checked >= 10


@pytest.mark.parametrize("W,weights", [(12, None), (16, (1, 3)), (20, (1, 2))])
def test_banded_against_oracle(W, weights):
    print("case fused banded W", W, flush=True)
    p = problem.banded(R=70, N=400, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=weights, seed=W)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=p.w)
    assert plan.feasible
    want = oracle.run_plan_c(plan, "oracle")
    assert fused.run_program(plan) == want
    assert kernel.run_plan(plan, executor="per-op")[0] == want


def test_full_size_equals_per_op():
    """2^25-entry banded plan (the headline workload's size): fused and per-op executors agree."""
    print("case fused full size W 25", flush=True)
    p = problem.banded(R=120, N=600, kmin=2, kmax=6, W=25, nobs=1, rho=0.1, seed=1)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    assert plan.feasible and plan.peak_bits >= 24
    prog = fused.program_for(plan)
    assert max(int(r["b_in"]) for r in prog.arr[:prog.npasses]) >= 24
    assert fused.run_program(plan) == kernel.run_plan(plan, executor="per-op")[0]


def test_weighted_overflow():
    """Every column of weight 200: any logical costs at least 400, saturates at 255 and raises OverflowError, as
    the per-op path does."""
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=5)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=[200] * p.ncols)
    with pytest.raises(OverflowError):
        oracle.run_plan_numpy(plan)
    prog = fused.compile_program(plan, min_bits=2)
    assert prog.passes and fused.simulate(prog, plan) == 255
    with pytest.raises(OverflowError):
        fused.run_program(plan, prog)


def test_bad_descriptor_rejected_on_host():
    """A descriptor whose addresses leave its table is refused before any pass kernel is launched."""
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=5)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    prog = fused.compile_program(plan, min_bits=4)
    segs, arr = fused.pack(prog)
    arr = arr.copy()
    assert arr[0]["nfree"] >= 1
    arr[0]["nruns"][0] = 1
    arr[0]["runs"][0, 0] = 0 | 1 << 8 | 38 << 16
    L = _lib.lib()
    ops = plan.encode()
    T0 = _lib.device_empty(1 << plan.peak_bits)
    T1 = _lib.device_empty(1 << plan.peak_bits)
    v = ctypes.c_int(-1)
    rc = L.tdp_run_program(ops.ctypes.data_as(ctypes.c_void_p), len(ops), segs.ctypes.data_as(ctypes.c_void_p),
                           len(segs), arr.ctypes.data_as(ctypes.c_void_p), len(arr), plan.peak_bits,
                           _lib.dptr(T0), _lib.dptr(T1), ctypes.byref(v), _lib.stream_handle())
    assert rc == 1    # TDP_EINVAL
    assert b"outside" in L.tdp_last_error()
    _lib.torch().cuda.synchronize()


def _cols(plan, order, used):
    ordinal = np.cumsum(plan.encode()[:, 0] == 1) - 1
    return sorted(order[plan.columns[int(ordinal[i])]] for i in used)


def test_fused_witness_random_against_brute_force():
    """Witness-mode program + one-wave traceback: value == brute force and a valid minimum solution, on the
    reference's random generator (half with costs 1-3)."""
    from tropidist_amd.problem import Problem, verify_witness
    print("case fused witness random", flush=True)
    rng = random.Random(5)
    checked = 0
    for t in range(200):
        m, k, h, l = random_problem(rng)
        w = [rng.randint(1, 3) for _ in h] if t % 2 else None
        order = list(range(len(h)))
        rng.shuffle(order)
        plan = trellis.build_plan(h, l, 1, order, m, weights=w)
        if not plan.feasible:
            continue
        v, used = fused.witness(plan)
        assert v == oracle.brute(h, l, 1, w), t
        assert verify_witness(Problem("r", m, k, h, l, w=w), _cols(plan, order, used), v, 0), t
        assert (v, used) == fused.witness_cpu(plan), t
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("W,weights", [(12, (1, 3)), (18, None), (22, (1, 3))])
def test_fused_witness_banded(W, weights):
    from tropidist_amd import auto
    from tropidist_amd.problem import verify_witness
    print("case fused witness banded W", W, flush=True)
    p = problem.banded(R=100, N=600, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=weights, seed=100 + W)
    order = list(range(p.ncols))
    plan = trellis.build_plan(p.h, p.l, 1, order, p.nrows, weights=p.w)
    assert plan.feasible
    want = kernel.run_plan(plan, executor="per-op")[0]
    v, used = fused.witness(plan)
    cols = _cols(plan, order, used)
    assert v == want and verify_witness(p, cols, v, 0)
    v2, cols2 = auto.witness_plan(plan, order, executor="per-op")
    assert v2 == v and verify_witness(p, cols2, v2, 0)


def test_fused_witness_recorded_instances():
    """auto.witness_plan (fused by default) on the recorded audit instances up to 2^22 entries: the recorded
    distance and a checked witness."""
    from tropidist_amd import auto
    from tropidist_amd.problem import Problem, verify_witness
    checked = 0
    for r in dense_fixtures():
        if r["distances"] is None or r["peak_bits"] > 22:
            continue
        p = Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"])
        for q, d in zip(r["plans"], r["distances"]):
            if d is None:
                continue
            print("case fused witness fixture", r["name"], q["observable"], flush=True)
            plan = trellis.build_plan(r["h"], r["l"], 1 << q["observable"], r["order"], r["nrows"])
            v, cols = auto.witness_plan(plan, r["order"])
            assert v == d and verify_witness(p, cols, v, q["observable"]), (r["name"], q["observable"])
            checked += 1
    assert checked >= 10


def test_fused_witness_overflow():
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=5)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=[200] * p.ncols)
    with pytest.raises(OverflowError):
        fused.witness(plan)


def test_run_many_equals_single_runs():
    """tdp_run_programs on a mix of value and witness runs of different widths (the rounds interleave their
    passes and per-op steps): every value equals the single-program executor's, every witness is valid."""
    from tropidist_amd.problem import verify_witness
    print("case run_many mixed", flush=True)
    rng = random.Random(77)
    plans, probs, orders = [], [], []
    for t in range(30):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        plan = trellis.build_plan(h, l, 1, order, m)
        if plan.feasible:
            plans.append(plan)
            probs.append(problem.Problem("r%d" % t, m, k, h, l))
            orders.append(order)
    for W, weights in ((10, None), (14, (1, 3)), (18, None), (20, (1, 2))):
        p = problem.banded(R=100, N=500, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=weights, seed=300 + W)
        order = list(range(p.ncols))
        plan = trellis.build_plan(p.h, p.l, 1, order, p.nrows, weights=p.w)
        assert plan.feasible
        plans.append(plan)
        probs.append(p)
        orders.append(order)
    witness = [i % 2 == 1 for i in range(len(plans))]
    got = fused.run_many(plans, witness)
    for i, (plan, p, order, w, g) in enumerate(zip(plans, probs, orders, witness, got)):
        want = fused.run_program(plan)
        if w:
            v, used = g
            assert v == want, i
            assert verify_witness(p, _cols(plan, order, used), v, 0), i
        else:
            assert g == want, i


def test_run_many_overflow():
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=5)
    bad = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=[200] * p.ncols)
    good = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    with pytest.raises(OverflowError):
        fused.run_many([good, bad])


def test_program_graph_replays_equal_run_many():
    """A captured schedule (tdp_programs_graph) replayed several times, with the program's own eager work and
    synchronisations between the replays: every replay gives run_many's values (widths 8-20, costs 1-3)."""
    from tropidist_amd import _lib
    print("case program graph replays", flush=True)
    plans = []
    for W, weights in ((8, None), (12, (1, 3)), (16, None), (20, (1, 2))):
        p = problem.banded(R=100, N=500, kmin=2, kmax=6, W=W, nobs=1, rho=0.1, weights=weights, seed=400 + W)
        plans.append(trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows, weights=p.w))
    want = fused.run_many(plans)
    g = fused.ProgramGraph(plans)
    assert g.nodes > len(plans)
    t = _lib.torch()
    for rep in range(4):
        g.launch()
        assert g.values() == want, rep
        assert fused.run_many(plans[:2]) == want[:2]          # eager work of the same program between replays
        t.cuda.synchronize()
    g.close()



def _run_small(plan, peak_bits=None, work_bytes=None):
    """tdp_run_plan_small through the C-ABI: (status, raw value)."""
    L = _lib.lib()
    ops = plan.encode()
    need = ctypes.c_uint64(0)
    _lib.check(L.tdp_plan_small_bytes(len(ops), ctypes.byref(need)), "tdp_plan_small_bytes")
    work = _lib.device_empty(need.value)
    value = ctypes.c_int(-1)
    rc = L.tdp_run_plan_small(ops.ctypes.data_as(ctypes.c_void_p), len(ops),
                              plan.peak_bits if peak_bits is None else peak_bits, work.data_ptr(),
                              need.value if work_bytes is None else work_bytes, ctypes.byref(value),
                              _lib.stream_handle())
    return rc, value.value


@pytest.mark.parametrize("W,weights", [(6, None), (8, None), (10, (1, 3)), (12, (1, 2)), (13, None)])
def test_small_plan_in_one_workgroup(W, weights):
    """Plans of at most 2^14 entries run whole in one workgroup's LDS: the value equals the oracle's and the
    per-op executor's, and run_plan's default takes that path for them."""
    print("case small plan W", W, flush=True)
    p = problem.banded(R=48, N=300, kmin=1, kmax=4, W=W, nobs=2, rho=0.1, weights=weights, seed=1000 + W)
    L = _lib.lib()
    for obs in range(p.nobs):
        plan = trellis.build_plan(p.h, p.l, 1 << obs, list(range(p.ncols)), p.nrows, weights=p.w)
        if not plan.feasible:
            continue
        assert plan.peak_bits <= L.tdp_plan_small_bits()
        want = oracle.run_plan_c(plan, "oracle")
        rc, v = _run_small(plan)
        assert rc == 0 and kernel.final_value(v) == want
        got, stats = kernel.run_plan(plan)
        assert got == want == kernel.run_plan(plan, executor="per-op")[0] and stats["executor"] == "small"


def test_small_plan_refusals():
    """The one-workgroup path refuses a plan wider than its LDS tables, a peak below the plan's, a workspace
    smaller than it needs, and a malformed op stream, before any launch; a wide plan asked for as "small" raises."""
    L = _lib.lib()
    p = problem.banded(R=48, N=300, kmin=1, kmax=4, W=8, nobs=1, rho=0.1, seed=1008)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    assert plan.feasible and plan.peak_bits >= 3
    assert _run_small(plan, peak_bits=L.tdp_plan_small_bits() + 1)[0] == _lib.TDP_EINVAL
    assert _run_small(plan, peak_bits=plan.peak_bits - 1)[0] == _lib.TDP_EINVAL
    assert _run_small(plan, work_bytes=16)[0] == _lib.TDP_EINVAL
    bad = plan.encode()[:-1].copy()                   # drops the final select: the stream ends with 2 entries
    need = ctypes.c_uint64(0)
    _lib.check(L.tdp_plan_small_bytes(len(bad), ctypes.byref(need)), "tdp_plan_small_bytes")
    work = _lib.device_empty(need.value)
    value = ctypes.c_int(-1)
    assert L.tdp_run_plan_small(bad.ctypes.data_as(ctypes.c_void_p), len(bad), plan.peak_bits, work.data_ptr(),
                                need.value, ctypes.byref(value), _lib.stream_handle()) == _lib.TDP_EINVAL
    wide = problem.banded(R=80, N=500, kmin=2, kmax=5, W=16, nobs=1, rho=0.1, seed=77)
    wplan = trellis.build_plan(wide.h, wide.l, 1, list(range(wide.ncols)), wide.nrows)
    if wplan.peak_bits > L.tdp_plan_small_bits():
        with pytest.raises(ValueError):
            kernel.run_plan(wplan, executor="small")
