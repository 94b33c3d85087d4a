"""GPU parity of the dense trellis DP: every libtdp kernel and the plan executor against the oracle, bit for bit."""
import ctypes
import random

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures, random_problem
from tropidist_amd import _lib, kernel, problem, trellis

pytestmark = pytest.mark.gpu


def to_dev(a):
    t = _lib.torch()
    return t.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(t):
    _lib.torch().cuda.synchronize()
    return t.cpu().numpy()


def S():
    return _lib.stream_handle()


def rand_table(rng, n, inf_frac=0.3):
    T = rng.integers(0, 256, size=n, dtype=np.uint8)
    T[rng.random(n) < inf_frac] = 255
    return T


@pytest.mark.parametrize("bits", [0, 1, 2, 3, 4, 5, 6, 8, 11, 16, 20])
def test_update_and_expand_match_oracle(bits):
    print("case update/expand bits", bits, flush=True)
    L, O = _lib.lib(), oracle.olib()
    rng = np.random.default_rng(100 + bits)
    n = 1 << bits
    masks = list(range(1, n)) if n <= 32 else sorted({int(x) for x in rng.integers(1, n, size=24)}
                                                     | {1, 15, 16, 17, n - 1, n >> 1, (n >> 1) | 5})
    for h in masks:
        for w in (1, 2, 3, 255):
            T = rand_table(rng, n)
            want = T.copy()
            O.or_dp_update_w(oracle.ptr(want), n, h, w)
            d = to_dev(T)
            _lib.check(L.tdp_update_w(_lib.dptr(d), n, h, w, S()), "update_w")
            assert np.array_equal(to_host(d), want), (n, h, w)
    for hl in [0] + masks[:12]:
        for w in (1, 3):
            T = rand_table(rng, 2 * n)
            want = T.copy()
            O.or_dp_expand_update_w(oracle.ptr(want), n, hl, w)
            d = to_dev(T)
            _lib.check(L.tdp_expand_update(_lib.dptr(d), n, hl, w, S()), "expand")
            assert np.array_equal(to_host(d), want), (n, hl, w)


@pytest.mark.parametrize("bits", [1, 2, 4, 5, 6, 9, 17])
def test_select_fill_hist_match_oracle(bits):
    print("case select/fill/hist bits", bits, flush=True)
    L, O = _lib.lib(), oracle.olib()
    rng = np.random.default_rng(200 + bits)
    n = 1 << bits
    for pos in range(bits):
        for target in (0, 1):
            T = rand_table(rng, n)
            want = T.copy()
            O.or_dp_select(oracle.ptr(want), n, pos, target)
            src, dst = to_dev(T), to_dev(np.zeros(n, np.uint8))
            nout = ctypes.c_uint64(0)
            _lib.check(L.tdp_select(_lib.dptr(dst), _lib.dptr(src), n, pos, target, ctypes.byref(nout), S()), "sel")
            assert nout.value == n // 2
            assert np.array_equal(to_host(dst)[:n // 2], want[:n // 2]), (n, pos, target)
    d = to_dev(np.zeros(n, np.uint8))
    _lib.check(L.tdp_fill(_lib.dptr(d), n, 255, S()), "fill")
    assert (to_host(d) == 255).all()
    T = rand_table(rng, n)
    want = np.zeros(256, np.uint64)
    O.or_dp_hist(oracle.ptr(T), n, oracle.ptr(want))
    hist = to_dev(np.full(256, 7, np.int64))
    _lib.check(L.tdp_hist(_lib.dptr(to_dev(T)), n, _lib.dptr(hist), S()), "hist")
    assert np.array_equal(to_host(hist).astype(np.uint64), want)


def test_run_plan_random_problems_against_brute_force():
    print("case run_plan random seed 1", flush=True)
    rng = random.Random(1)
    feasible = 0
    for t in range(1500):
        m, k, h, l = random_problem(rng)
        order = list(range(len(h)))
        rng.shuffle(order)
        obs_mask = rng.randrange(1, 1 << k)
        plan = trellis.build_plan(h, l, obs_mask, order, m)
        got, stats = kernel.run_plan(plan)
        assert got == oracle.brute(h, l, obs_mask), t
        feasible += got is not None
    assert feasible > 900


def test_run_plan_weighted_against_brute_force():
    print("case run_plan weighted seed 11", flush=True)
    rng = random.Random(11)
    for t in range(400):
        m, k, h, l = random_problem(rng)
        w = [rng.randint(1, 3) for _ in h]
        order = list(range(len(h)))
        rng.shuffle(order)
        obs_mask = rng.randrange(1, 1 << k)
        plan = trellis.build_plan(h, l, obs_mask, order, m, weights=w)
        assert kernel.run_plan(plan)[0] == oracle.brute(h, l, obs_mask, w), t


def test_chain_boundary_on_gpu():
    print("case chain boundary", flush=True)
    for n in (254, 255, 256):
        p = problem.chain(n)
        plan = trellis.build_plan(p.h, p.l, 1, list(range(n)), p.nrows)
        if n < 255:
            assert kernel.run_plan(plan)[0] == n
        else:
            with pytest.raises(OverflowError):
                kernel.run_plan(plan)
    p = problem.Problem("no logical", 1, 1, [1, 1], [1, 1])
    assert kernel.run_plan(trellis.build_plan(p.h, p.l, 1, [0, 1], p.nrows))[0] is None


@pytest.mark.parametrize("executor", kernel.EXECUTORS)
def test_recorded_distances_on_gpu(executor):
    """Every observable of the fixtures with tables up to 2^28 entries: the recorded distance."""
    for r in dense_fixtures():
        if r["distances"] is None or r["peak_bits"] > 28:
            continue
        print("case fixture", executor, r["name"], r["variant"], "peak", r["peak_bits"], flush=True)
        for p, d in zip(r["plans"], r["distances"]):
            plan = trellis.build_plan(r["h"], r["l"], 1 << p["observable"], r["order"], r["nrows"])
            assert kernel.run_plan(plan, executor=executor)[0] == d, (r["name"], p["observable"])


@pytest.mark.parametrize("executor", kernel.EXECUTORS)
def test_banded_width20_equals_oracle(executor):
    """A weighted banded model at width 20 (2^20-entry tables): GPU == oracle C on every observable."""
    print("case banded W20", executor, flush=True)
    p = problem.banded(R=120, N=700, kmin=2, kmax=6, W=20, nobs=2, rho=0.05, weights=(1, 3), seed=2020)
    order = list(range(p.ncols))
    for obs in range(p.nobs):
        plan = trellis.build_plan(p.h, p.l, 1 << obs, order, p.nrows, weights=p.w)
        if not plan.feasible:
            continue
        assert plan.peak_bits >= 18
        assert kernel.run_plan(plan, executor=executor)[0] == oracle.run_plan_c(plan, "oracle")


class _RefStylePlan:
    """Only the attributes the reference's Plan has (cdist/trellis.py:125-144): ops as tuples, peak_bits, feasible."""

    def __init__(self, plan):
        self.ops = [tuple(op) for op in plan.ops]
        self.peak_bits = plan.peak_bits
        self.feasible = plan.feasible


@pytest.mark.parametrize("executor", kernel.EXECUTORS)
def test_run_plan_accepts_reference_style_plans(executor):
    """The drop-in surface: a plan object built like the reference's runs unchanged through run_plan."""
    p = problem.banded(R=80, N=500, kmin=2, kmax=5, W=14, nobs=1, rho=0.1, seed=77)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    want = oracle.run_plan_c(plan, "oracle")
    got, stats = kernel.run_plan(_RefStylePlan(plan), executor=executor)
    assert got == want and stats["threads"] == stats["cus"] > 0
