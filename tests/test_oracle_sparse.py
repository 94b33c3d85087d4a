"""Sparse threshold DP: the oracle's step / merge / join equal the reference's compiled sparse_dp.c, and the
product's host logic (Trellis, Sweep, query, distance, schedules) driven by the oracle tables equals brute force
(the patterns of tests/test_sparse_dp.py and test_false_bounds.py:205-227)."""
import random

import numpy as np
import pytest

import oracle
from sparse_helpers import brute, check_witness, random_problem
from tropidist_amd import problem, sparse_dp

needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")
BACKENDS = [oracle.OracleTables] + ([oracle.RefTables] if oracle.have_ref() else [])


class _Tr:
    nslots = 20

    def bound_data(self, forward):
        return []


@needs_ref
def test_step_and_join_equal_reference_c():
    rng = np.random.default_rng(5)
    a, b = oracle.OracleTables(_Tr()), oracle.RefTables(_Tr())
    for trial in range(400):
        bits = int(rng.integers(1, 20))
        n = int(rng.integers(0, 300))
        keys = np.unique(rng.integers(0, 1 << bits, size=n).astype(np.uint64))
        vals = rng.integers(0, 9, size=keys.size).astype(np.uint8)
        tab = a.from_arrays(keys, vals)
        upd = int(rng.integers(0, 1 << bits))
        sm = int(rng.integers(0, 1 << bits)) & int(rng.integers(0, 1 << bits))
        sb = sm & int(rng.integers(0, 1 << bits))
        cap, kb = int(rng.integers(0, 9)), int(rng.integers(0, 10))
        ncls = int(rng.integers(0, 3))
        hd = None
        if ncls:
            masks = rng.integers(0, 1 << bits, size=ncls).astype(np.uint64)
            ws = rng.integers(1, 720720, size=ncls).astype(np.int64)
            hd = (masks, ws, ncls, int(rng.integers(0, 720720)))
        tm = int(rng.integers(0, 1 << bits))
        K = int(rng.integers(0, 12))
        ra = a.step(tab, upd, sm, sb, cap, K, tm, hd, kb, True)
        rb = b.step(tab, upd, sm, sb, cap, K, tm, hd, kb, True)
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1]), trial
        bkeys = np.unique(rng.integers(0, 1 << bits, size=60).astype(np.uint64))
        bt = a.from_arrays(bkeys, rng.integers(0, 9, size=bkeys.size).astype(np.uint8))
        tmask = 0 if rng.random() < 0.5 else 1 << int(rng.integers(0, bits))
        assert a.join(ra[1], bt, tmask) == b.join(ra[1], bt, tmask), trial


def test_random_queries_all_thresholds_and_splits():
    """tests/test_sparse_dp.py:49-74 (seed 3, 60 problems): every K, forward and every split, with and without the
    lookahead, through the oracle tables."""
    rng = random.Random(3)
    checked = 0
    for _ in range(60):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            truth = brute(p, obs)
            assert tr.feasible == (truth is not None)
            if not tr.feasible:
                continue
            for K in range(1, p.ncols + 2):
                want = truth if truth <= K else None
                for mode, split in [("forward", None)] + [("mitm", a) for a in range(1, K + 1)]:
                    for use_h in (False, True):
                        be = BACKENDS[checked % len(BACKENDS)](tr)
                        r = sparse_dp.query(tr, K, mode=mode, split=split, use_h=use_h, backend=be,
                                            every=rng.choice([1, 2, 5]))
                        assert r["value"] == want, (mode, split, use_h, K, r["value"], want)
                        if want is not None:
                            check_witness(p, r["witness"], want, obs)
                        checked += 1
    assert checked > 1000


def test_split_schedules():
    """tests/test_sparse_dp.py:77-107 (seed 9): falling splits and the chosen schedule; a rising split is
    refused."""
    rng = random.Random(9)
    checked = 0
    for _ in range(50):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            if not tr.feasible:
                continue
            truth = brute(p, obs)
            for K in range(1, p.ncols + 2):
                want = truth if truth <= K else None
                splits = [sorted((rng.randint(1, K) for _ in range(tr.N + 1)), reverse=True) for _ in range(3)]
                for split in splits + ["schedule"]:
                    r = sparse_dp.query(tr, K, mode="mitm", split=split, backend=oracle.OracleTables(tr),
                                        use_h=rng.random() < 0.5, every=rng.choice([1, 3]))
                    assert r["value"] == want, (split, K, r["value"], want)
                    if want is not None:
                        check_witness(p, r["witness"], want, obs)
                    checked += 1
    assert checked > 500
    tr = sparse_dp.Trellis.for_query(problem.chain(6), 0, list(range(6)))
    with pytest.raises(AssertionError):
        sparse_dp.query(tr, 3, mode="mitm", split=list(range(1, tr.N + 2)), backend=oracle.OracleTables(tr))


def test_distance_driver_matches_brute_force():
    rng = random.Random(5)
    for _ in range(80):
        p = random_problem(rng)
        truth = [brute(p, i) for i in range(p.nobs)]
        truth = min((t for t in truth if t is not None), default=None)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for mode in ("forward", "mitm"):
            tr0 = sparse_dp.Trellis.for_query(p, 0, order)
            be = oracle.OracleTables(tr0) if tr0.feasible else oracle.OracleTables(_Tr())
            r = sparse_dp.distance(p, order=order, mode=mode, kmax=p.ncols + 1, backend=be)
            assert r["distance"] == truth
            if truth is not None:
                check_witness(p, r["witness"], truth)


def test_sparse_8bit_boundary():
    """tests/test_false_bounds.py:205-227: a forward sweep keeps weights <= K: 254 exact, 255 and 256 refused; a
    meet in the middle is exact at 255 and 256 and refuses 510."""
    for n in (254, 255, 256):
        p = problem.chain(n)
        tr = sparse_dp.Trellis.for_query(p, 0, list(range(n)))
        if n == 254:
            r = sparse_dp.query(tr, n, mode="forward", backend=oracle.OracleTables(tr), use_h=False)
            assert r["value"] == n and r["witness"] == list(range(n))
        else:
            with pytest.raises(ValueError):
                sparse_dp.query(tr, n, mode="forward", backend=oracle.OracleTables(tr), use_h=False)
        r = sparse_dp.query(tr, n, mode="mitm", backend=oracle.OracleTables(tr), use_h=False)
        assert r["value"] == n and r["witness"] == list(range(n))
    tr = sparse_dp.Trellis.for_query(problem.chain(510), 0, list(range(510)))
    with pytest.raises(ValueError):
        sparse_dp.query(tr, 510, mode="mitm", backend=oracle.OracleTables(tr), use_h=False, witness=False)


class _OracleEngine:
    """The engine's query interface over the oracle's CPU tables: query `rid` is sparse_dp.query at the threshold
    it was added with (the words argument carries the Trellis)."""

    def __init__(self):
        self.q, self.released = {}, set()
        self.max_states = None

    def add_words(self, tr, tm, N, K):
        rid = len(self.q)
        self.q[rid] = sparse_dp.query(tr, K, mode="mitm", use_h=False, witness=True, backend=oracle.OracleTables(tr))
        return rid

    def result(self, rid):
        return 1, self.q[rid]["value"], {"peak_states": self.q[rid]["peak_states"]}

    def witness(self, rid, order, value):
        return sorted(self.q[rid]["witness"])

    def release(self, rid):
        assert rid not in self.released
        self.released.add(rid)

    def cancel(self, rid):
        pass


def test_request_deepening_equals_brute_force():
    """sparse_batch.Request deepens through the threshold ladder; with the oracle answering its queries, its
    distance and observable are the least weight <= K over the observables, as brute force gives it."""
    from tropidist_amd import sparse_batch
    rng = random.Random(11)
    checked = 0
    for _ in range(60):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        trs = [(i, sparse_dp.Trellis.for_query(p, i, order), None) for i in range(p.nobs)]
        trs = [t for t in trs if t[1].feasible]
        if not trs:
            continue
        truth = {i: brute(p, i) for i, _, _ in trs}
        for K in range(1, p.ncols + 1):
            req = sparse_batch.Request(p, K, order=order)
            req.trs = trs
            eng = _OracleEngine()
            while True:
                req.admit(eng)
                for rid in req.ids:
                    req.collect(eng, rid)
                req.finish(eng)
                if not req.pending:
                    break
            assert req.error is None, req.error
            within = {i: t for i, t in truth.items() if t is not None and t <= K}
            best = min(within.values(), default=None)
            assert req.record["distance"] == best
            if best is not None:
                assert req.record["observable"] == min(i for i, t in within.items() if t == best)
                assert req.record["witness_ok"]
            assert len(eng.released) == len(eng.q)        # every query released once
            checked += 1
    assert checked > 100


def _cut_tables(tr, direction, K, use_h, dense, be, dense_max=1 << 30):
    """{cut: sorted (key, value) pairs} of one whole sweep keeping weights <= K, and the sweep."""
    sw = sparse_dp.Sweep(tr, direction, K, K, use_h, backend=be, dense=dense, dense_max=dense_max)
    out = {}

    def keep(c, tab, changed, extra):
        keys, vals = sw.be.arrays(tab)
        out[c] = sorted(zip((int(k) for k in keys), (int(v) for v in vals)))

    sw.run(on_cut=keep)
    return out, sw


def test_dense_fallback_keeps_the_sparse_tables():
    """The dense form changes the representation only: a sweep with the fallback (every table dense from the first
    step at 1e-9, switching back and forth at 0.25) holds, at every cut, exactly the states and values of the plain
    sparse sweep, which test_step_and_join_equal_reference_c and the brute-force tests pin.  dense_max = 8 makes
    the wider cuts switch back to sparse tables and the narrower ones dense again."""
    rng = random.Random(11)
    dense_cuts = 0
    for _ in range(40):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            if not tr.feasible:
                continue
            for direction in (+1, -1):
                K = rng.randint(1, p.ncols)
                use_h = rng.random() < 0.5
                plain, _ = _cut_tables(tr, direction, K, use_h, None, oracle.OracleTables(tr))
                for dense, dmax in ((1e-9, 1 << 30), (0.25, 1 << 30), (1e-9, 8)):
                    got, sw = _cut_tables(tr, direction, K, use_h, dense, oracle.OracleTables(tr), dmax)
                    assert got == plain, (direction, K, use_h, dense, dmax)
                    dense_cuts += sw.dense_cuts
    assert dense_cuts > 100


def test_dense_fallback_queries_against_brute_force():
    """tests/test_sparse_dp.py:49-74 and 146-155 with the fallback: forward sweeps dense from the first step, and
    meets in the middle whose backward sweep goes dense, answer every threshold as brute force does."""
    rng = random.Random(3)
    checked = dense_cuts = 0
    for _ in range(40):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            if not tr.feasible:
                continue
            truth = brute(p, obs)
            for K in range(1, p.ncols + 2):
                want = truth if truth <= K else None
                for mode, split, dense in (("forward", None, 1e-9), ("mitm", K // 2 + 1, 0.25), ("mitm", 1, 1e-9)):
                    r = sparse_dp.query(tr, K, mode=mode, split=split, use_h=rng.random() < 0.5, dense=dense,
                                        backend=oracle.OracleTables(tr), every=rng.choice([1, 2, 5]))
                    assert r["value"] == want, (mode, split, dense, K, r["value"], want)
                    if want is not None:
                        check_witness(p, r["witness"], want, obs)
                    checked += 1
                    dense_cuts += r["dense_cuts"]
    assert checked > 500 and dense_cuts > 1000


def test_dense_fallback_at_the_8bit_boundary():
    """tests/test_false_bounds.py:213-225 with dense=0.25: K = 254 forward and 256 in the middle stay exact."""
    for n, mode in ((254, "forward"), (256, "mitm")):
        tr = sparse_dp.Trellis.for_query(problem.chain(n), 0, list(range(n)))
        r = sparse_dp.query(tr, n, mode=mode, backend=oracle.OracleTables(tr), use_h=False, dense=0.25)
        assert r["value"] == n and r["witness"] == list(range(n)) and r["dense_cuts"] > 0
