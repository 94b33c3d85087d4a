"""GPU parity of the sparse threshold DP with meet in the middle: device tables (HipTables) against brute force and
against the oracle tables state for state (identical per-cut state counts), on random problems, the reference's
known codes and the 8-bit boundary."""
import random

import pytest

import oracle
from helpers import dense_fixtures
from sparse_helpers import brute, check_witness, random_problem
from tropidist_amd import problem, sparse_dp
from tropidist_amd.problem import Problem

pytestmark = pytest.mark.gpu


def test_random_queries_against_brute_force():
    print("case sparse random seed 3", flush=True)
    rng = random.Random(3)
    checked = 0
    for _ in range(40):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            truth = brute(p, obs)
            if not tr.feasible:
                continue
            for K in range(1, p.ncols + 2):
                want = truth if truth <= K else None
                for mode, split in [("forward", None)] + [("mitm", a) for a in range(1, K + 1, 2)]:
                    use_h = rng.random() < 0.5
                    r = sparse_dp.query(tr, K, mode=mode, split=split, use_h=use_h, backend="hip",
                                        every=rng.choice([1, 2, 5]))
                    assert r["value"] == want, (mode, split, use_h, K, r["value"], want)
                    if want is not None:
                        check_witness(p, r["witness"], want, obs)
                    checked += 1
    assert checked > 300


def _same_sweeps(tr, K, split, use_h):
    a = sparse_dp.query(tr, K, mode="mitm", split=split, use_h=use_h, backend="hip")
    b = sparse_dp.query(tr, K, mode="mitm", split=split, use_h=use_h, backend=oracle.OracleTables(tr))
    assert a["value"] == b["value"]
    for key in ("forward_counts", "backward_counts", "changed_counts"):
        assert a[key] == b[key], key
    return a


def test_state_counts_equal_oracle_on_recorded_codes():
    """code_bb72 (d = 6) and star_42 (d = 4): every cut's table and changed-state count equals the oracle's."""
    for r in dense_fixtures():
        if r["name"] not in ("code_bb72", "star_42"):
            continue
        p = Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"])
        d = min(x for x in r["distances"] if x is not None)
        for obs in range(min(2, p.nobs)):
            tr = sparse_dp.Trellis.for_query(p, obs, r["order"])
            if not tr.feasible or tr.nslots > 56:
                continue
            print("case sparse counts", r["name"], "obs", obs, "slots", tr.nslots, flush=True)
            for K, use_h in ((d, False), (d - 1, True)):
                rec = _same_sweeps(tr, K, K // 2 + 1, use_h)
                if rec["value"] is not None:
                    check_witness(p, rec["witness"], rec["value"], obs)


def test_known_code_distances():
    """bb72 = 6 and bb90 = 10 by iterative deepening from d - 1 (tests/test_sparse_dp.py:158-167)."""
    for r in dense_fixtures():
        if r["name"] not in ("code_bb72", "code_bb90"):
            continue
        d = min(x for x in r["distances"] if x is not None)
        print("case sparse distance", r["name"], d, flush=True)
        p = Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"])
        out = sparse_dp.distance(p, order=r["order"], mode="mitm", backend="hip", start=d - 1, use_h=False)
        assert out["distance"] == d and out["lower_bound"] == d
        check_witness(p, out["witness"], d)


def test_sparse_8bit_boundary_on_gpu():
    print("case sparse chain boundary", flush=True)
    for n in (254, 256):
        tr = sparse_dp.Trellis.for_query(problem.chain(n), 0, list(range(n)))
        if n == 254:
            r = sparse_dp.query(tr, n, mode="forward", backend="hip", use_h=False)
            assert r["value"] == n and r["witness"] == list(range(n))
        else:
            with pytest.raises(ValueError):
                sparse_dp.query(tr, n, mode="forward", backend="hip", use_h=False)
        r = sparse_dp.query(tr, n, mode="mitm", backend="hip", use_h=False)
        assert r["value"] == n and r["witness"] == list(range(n))


def test_wide_quasi_cyclic_code_beyond_64_slots():
    """A two-block QC code whose trellis needs more row slots than a 64-bit key: the 120-bit device keys run it;
    the witness is a valid solution and the schedule choice agrees with a fixed split."""
    p = problem.quasi_cyclic(6, 6, seed=3)
    order = list(range(p.ncols))
    tr = sparse_dp.Trellis.for_query(p, 0, order)
    print("case sparse qc 6x6 slots", tr.nslots, flush=True)
    a = sparse_dp.query(tr, 8, mode="mitm", split="schedule", use_h=False, backend="hip")
    b = sparse_dp.query(tr, 8, mode="mitm", split=5, use_h=True, backend="hip")
    assert a["value"] == b["value"]
    if a["value"] is not None:
        check_witness(p, a["witness"], a["value"], 0)


def test_sp_hist_counts_values():
    """tdp_sp_hist (sp_hist, cdist/sparse_dp.c:238-242) on device tables, keys in both state words: the value
    histogram equals numpy's count of the stored values."""
    import numpy as np
    print("case sparse hist", flush=True)
    rng = random.Random(8)
    tr = None
    while tr is None or not tr.feasible:
        p = random_problem(rng)
        tr = sparse_dp.Trellis.for_query(p, 0, list(range(p.ncols)))
    be = sparse_dp.HipTables(tr)
    for n in (1, 7, 1000, 70000):
        keys = set()
        while len(keys) < n:
            keys.add(rng.getrandbits(100))
        keys = sorted(keys)
        vals = np.array([rng.randrange(0, 255) for _ in range(n)], dtype=np.uint8)
        tab = be.from_arrays(keys, vals)
        assert list(be.hist(tab)) == np.bincount(vals, minlength=256).tolist()
