"""The sparse path with more than 64 row slots (keys that use the upper state word), pinned by recorded distances.

code_bb144 ([[144,12,12]], 72 checks) has distance 12 on every observable (recorded by the reference's dense audit,
results/dense_audit_20261021).  The distance does not depend on the column order, but the number of row slots does:
in a shuffled order the trellis needs 66 slots, so every key spans both words of a device state.  Threshold 11 must
give no solution (a certified lower bound through the upper key word) and threshold 12 the recorded 12, with a
witness checked against H and L.  On the quasi-cyclic l = m = 12 code (77 slots) the one-query path and the batched
engine must agree at threshold 12."""
import random
import time

import pytest

from helpers import dense_fixtures
from sparse_helpers import check_witness
from tropidist_amd import ordering, problem, sparse_batch, sparse_dp, workloads
from tropidist_amd.problem import Problem

pytestmark = pytest.mark.gpu


def _bb144():
    for r in dense_fixtures():
        if r["name"] == "code_bb144":
            return Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"]), r
    raise AssertionError("code_bb144 fixture missing")


def test_bb144_shuffled_order_beyond_64_slots():
    p, r = _bb144()
    order = list(r["order"])
    random.Random(0).shuffle(order)
    tr = sparse_dp.Trellis.for_query(p, 0, order)
    assert tr.nslots > 64, tr.nslots
    assert r["distances"][0] == 12
    for K, want in ((11, None), (12, 12)):
        print("case bb144 shuffled slots", tr.nslots, "K", K, flush=True)
        t = time.time()
        rec = sparse_dp.query(tr, K, mode="mitm", use_h=False, backend="hip")
        print("  value", rec["value"], "peak states", rec["peak_states"], "%.2f s" % (time.time() - t), flush=True)
        assert rec["value"] == want
        if want is not None:
            check_witness(p, rec["witness"], want, 0)


def _single_vs_engine(p, tr, K):
    one = sparse_dp.query(tr, K, mode="mitm", use_h=False, backend="hip")
    _lib_torch().cuda.empty_cache()
    eng = sparse_batch.Engine(arena_bytes=96 << 30, max_states=0)
    try:
        rid = eng.add(tr, K)
        while eng.step():
            pass
        status, value, stats = eng.result(rid)
        assert status == sparse_batch.DONE, status
        assert value == one["value"], (value, one["value"])
        if value is not None:
            cols = eng.witness(rid, tr, value)
            check_witness(p, cols, value, 0)
            check_witness(p, one["witness"], value, 0)
        eng.release(rid)
    finally:
        eng.close()
    return one["value"]


def _lib_torch():
    from tropidist_amd import _lib
    return _lib.torch()


@pytest.mark.parametrize("K", [10, 11])
def test_qc12_single_query_equals_engine(K):
    """The l = m = 12 code in the greedy order (77 slots); the engine keeps every table, so thresholds whose
    sweeps fit its arena."""
    p = workloads.sparse_qc()
    order = ordering.greedy_straddle_order(p.h, p.nrows)
    tr = sparse_dp.Trellis.for_query(p, 0, order)
    assert tr.nslots > 64
    print("case qc12 single vs engine, slots", tr.nslots, "K", K, flush=True)
    _single_vs_engine(p, tr, K)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_qc9_shuffled_single_query_equals_engine(seed):
    """l = m = 9 codes (162 columns) in shuffled orders: 72-77 slots."""
    p = problem.quasi_cyclic(9, 9, seed=seed)
    order = list(range(p.ncols))
    random.Random(seed).shuffle(order)
    tr = sparse_dp.Trellis.for_query(p, 0, order)
    assert tr.nslots > 64
    for K in (6, 8):
        print("case qc9 seed", seed, "slots", tr.nslots, "K", K, flush=True)
        _single_vs_engine(p, tr, K)
