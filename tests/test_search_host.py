"""The weight-bounded search's host side and its checker, without a GPU.

* Searcher's incidence arrays (rows list their columns in increasing index: the branching order of the search) and
  exclusion pairs against their definitions.
* The checker: the reference's own gm_search (gmitm.c compiled from source into oracle/_ref), driven with the
  Searcher's arrays by oracle.gm_search_ref, gives the brute-force distance and every minimum solution on random
  problems (the check of tests/test_search.py in the reference), in both table modes, with and without exclusions,
  the covering bound, and the forced fingerprint collision of tests/test_collision.py."""
import random

import numpy as np
import pytest

import oracle
from tropidist_amd import search, tables
from tropidist_amd.problem import Problem, dedupe

needs_ref = pytest.mark.skipif(not oracle.have_gmitm(), reason="oracle/_ref/libgmitm.so not built")


def random_problem(rng, nobs=None, high_labels=False):
    m, n = rng.randint(2, 10), rng.randint(2, 16)
    k = nobs or rng.randint(1, 3)
    h = [sum(1 << i for i in range(m) if rng.random() < rng.choice([0.2, 0.35])) for _ in range(n)]
    if high_labels:
        l = [sum(1 << rng.randrange(16, k) for _ in range(rng.randint(0, 2))) for _ in range(n)]
    else:
        l = [sum(1 << i for i in range(k) if rng.random() < 0.3) for _ in range(n)]
    h, l, _ = dedupe(h, l)
    return Problem("rand", m, k, h, l, [()] * m) if h else None


def test_incidence_arrays():
    rng = random.Random(5)
    for _ in range(50):
        p = random_problem(rng)
        if p is None:
            continue
        S = search.Searcher(p)
        for r in range(p.nrows):
            cols = S.row_cols[S.row_start[r]:S.row_start[r + 1]].tolist()
            assert cols == [j for j in range(p.ncols) if p.h[j] >> r & 1]
        for j in range(p.ncols):
            rows = S.col_rows[S.col_start[j]:S.col_start[j + 1]].tolist()
            assert rows == [r for r in range(p.nrows) if p.h[j] >> r & 1]
            assert sum(int(S.col[j, w]) << (64 * w) for w in range(S.W)) == p.h[j]
        assert S.wmax == max(bin(x).count("1") for x in p.h)


def test_exclusion_pairs_match_their_definition():
    rng = random.Random(6)
    for _ in range(60):
        p = random_problem(rng)
        if p is None:
            continue
        S = search.Searcher(p)
        npairs = S.build_exclusions()
        want = [set() for _ in range(p.ncols)]
        index = {}
        for j in range(p.ncols):
            index[(p.h[j], p.l[j])] = j
        for x in range(p.ncols):
            for y in range(x + 1, p.ncols):
                if not p.h[x] & p.h[y]:
                    continue
                z = index.get((p.h[x] ^ p.h[y], p.l[x] ^ p.l[y]))
                if z is not None:
                    want[x] |= {y, z}
                    want[y] |= {x, z}
                    want[z] |= {x, y}
        got = [set(S.ex_cols[S.ex_start[j]:S.ex_start[j + 1]].tolist()) for j in range(p.ncols)]
        assert got == want
        assert npairs == sum(len(s) for s in want) // 2


def test_ledger_and_threshold_limits():
    search.set_ledger()
    with pytest.raises(NotImplementedError):
        search.set_ledger("/tmp/ledger")
    with pytest.raises(ValueError):
        search.set_ledger(part=(2, 2))
    with pytest.raises(ValueError):
        search.check_threshold(41)


def ref_tables(S, p, k, labelfree, seed=12345):
    rh = tables.row_hashes(p.nrows, seed)
    zh, zl = tables.column_hashes(p.h, rh)
    lg = np.zeros(p.ncols, dtype=np.uint64) if labelfree else np.array(p.l, dtype=np.uint64)
    wide = int(p.nobs > 16 and not labelfree)
    return [(zh, zl, lg, k, tables.table_log2slots(p.ncols, k), wide)]


def ref_distance(p, kt, enum_all=True, exclusions=True, labelfree=False, cover_gate=0):
    S = search.Searcher(p, labelfree=labelfree)
    if exclusions:
        S.build_exclusions()
    tabs = ref_tables(S, p, kt, labelfree)
    for K in range(1, p.ncols + 1):
        r = oracle.gm_search_ref(S, K, tabs, 0, S.rank, S.anchors, use_ex=exclusions, enum_all=enum_all,
                                 task_depth=max(1, min(3, K - kt - 1)), cover_gate=cover_gate)
        assert r["status"] in (0, 1)
        if r["best"] is not None:
            return r["best"], sorted({s for s in r["solutions"] if len(s) == r["best"]})
    return None, []


@needs_ref
def test_checker_equals_brute_force():
    rng = random.Random(3)
    done = 0
    for _ in range(120):
        p = random_problem(rng)
        if p is None:
            continue
        want, wsols = oracle.brute_all(p.h, p.l)
        for kt in (1, 2, 3):
            for ex in (False, True):
                d, sols = ref_distance(p, kt, exclusions=ex, labelfree=(kt == 2))
                assert d == want
                if want is not None:
                    assert sols == sorted(wsols)
        done += 1
    assert done > 80


@needs_ref
def test_checker_wide_labels_and_cover():
    rng = random.Random(9)
    for _ in range(60):
        p = random_problem(rng, nobs=rng.randint(17, 64), high_labels=True)
        if p is None:
            continue
        want, wsols = oracle.brute_all(p.h, p.l)
        d, sols = ref_distance(p, 2, cover_gate=4)
        assert d == want
        if want is not None:
            assert sols == sorted(wsols)


@needs_ref
def test_checker_forced_collision():
    """gmitm.c completes every hit with the full budget: an all-colliding table must not hide the weight-2 solution
    (tests/test_collision.py of the reference)."""
    p = Problem("collide", 1, 2, [1, 1], [1, 2], [()])
    for labelfree in (False, True):
        S = search.Searcher(p, labelfree=labelfree)
        z = np.zeros(2, dtype=np.uint64)
        lg = z if labelfree else np.array(p.l, dtype=np.uint64)
        tabs = [(z, z, lg, 1, 10, 0)]
        r1 = oracle.gm_search_ref(S, 1, tabs, 0, S.rank, S.anchors, use_ex=False)
        r2 = oracle.gm_search_ref(S, 2, tabs, 0, S.rank, S.anchors, use_ex=False)
        assert r1["status"] == 0 and not r1["solutions"]
        assert r2["solutions"] == [(0, 1)]
