"""The weight-bounded search on the GPU (tdp_gm_search through tropidist_amd.search).

* Random problems: distance and every minimum solution equal brute force, for table depths 1-3, with and without
  exclusions, labelled and label-free, and with more than 16 observables (wide tables) -- tests/test_search.py of
  the reference.
* The forced fingerprint collision of the reference's tests/test_collision.py.
* Against the reference's own gm_search (oracle/_ref, gmitm.c compiled from source), threshold by threshold on
  structured codes: status, best weight, the solution set, nodes visited per depth, hits, false hits, tasks and
  cover/ascent prunes -- exactly -- with the covering bound, exclusions, kernel dominance sets, a pattern budget,
  relaxed tables and label-free mode.
* The time limit ends a run with status "timeout"."""
import random

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures
from test_search_host import random_problem, ref_tables
from tropidist_amd import problem, search, tables
from tropidist_amd.problem import Problem

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not oracle.have_gmitm(), reason="oracle/_ref/libgmitm.so not built")


@pytest.mark.parametrize("seed", [3, 4])
def test_random_problems_against_brute_force(seed):
    print("case search random seed", seed, flush=True)
    rng = random.Random(seed)
    done = 0
    for t in range(150):
        p = random_problem(rng)
        if p is None:
            continue
        want, wsols = oracle.brute_all(p.h, p.l)
        for kt in (1, 2, 3):
            for ex in (False, True):
                r = search.distance(p, k_table=kt, kmax=p.ncols, enum_all=True, exclusions=ex,
                                    labelfree=(t + kt) % 2 == 0, time_limit=60.0)
                assert r["distance"] == want, (t, kt, ex, want, r["distance"])
                if want is not None:
                    assert r["solutions"] == sorted(wsols), (t, kt, ex)
                    assert r["lower_bound"] == want
        done += 1
    assert done > 100


def test_wide_labels_against_brute_force():
    print("case search wide labels", flush=True)
    rng = random.Random(9)
    for t in range(120):
        p = random_problem(rng, nobs=rng.randint(17, 64), high_labels=True)
        if p is None:
            continue
        want, wsols = oracle.brute_all(p.h, p.l)
        for kt in (1, 2, 3):
            r = search.distance(p, k_table=kt, kmax=p.ncols, enum_all=True, time_limit=60.0)
            assert r["distance"] == want, (t, kt)
            if want is not None:
                assert r["solutions"] == sorted(wsols)


def test_forced_collision():
    print("case search forced collision", flush=True)
    p = Problem("collide", 1, 2, [1, 1], [1, 2], [()])
    real = tables.row_hashes
    for labelfree in (False, True):
        S = search.Searcher(p, labelfree=labelfree)
        tables.row_hashes = lambda nrows, seed: ([0] * nrows, [0] * nrows)
        try:
            tab = tables.Table(p, 1, labelfree=labelfree)
        finally:
            tables.row_hashes = real
        S.add_table(tab, exact=True)
        r1, r2 = S.search(1), S.search(2)
        tab.free()
        assert r1["status"] == "exhausted" and not r1["solutions"]
        assert r2["solutions"] and sorted(r2["solutions"][0]) == [0, 1]


def _codes():
    """Recorded codes with their distances: deep enough trees (up to 8e5 nodes) for the counts to mean something."""
    names = ("code_bb72", "code_bb90", "star_42", "routing_54", "surface_d5", "bb72_r6", "sdbb_64", "walkingcat_54")
    out = []
    for r in dense_fixtures():
        if r["name"] in names and r["name"] not in [p.name for p, _ in out]:
            out.append((Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"]),
                        min(x for x in r["distances"] if x is not None)))
    return out


def _compare(S, p, K, kt, labelfree, *, cover=0, exclusions=True, enum_all=True, kernel=None, pattern=None,
             rank=None, anchors=None, relaxed=()):
    """One threshold on the device and in the reference; the records must agree exactly."""
    tab = tables.Table(p, kt, labelfree=labelfree)
    try:
        S.tables, S.main = [], -1
        ref_tabs = []
        for rt, rzh, rzl, rlg in relaxed:
            S.add_table(rt, rzh, rzl)
            ref_tabs.append((rzh, rzl, rlg, rt.k, rt.log2slots, int(rt.wide)))
        S.add_table(tab, exact=True)
        ref_tabs += ref_tables(S, p, kt, labelfree)
        rk = S.rank if rank is None else np.asarray(rank, dtype=np.int32)
        an = S.anchors if anchors is None else np.asarray(anchors, dtype=np.int32)
        td = max(1, min(4 if cover else 3, K - kt - 1))
        got = S.search(K, enum_all=enum_all, task_depth=td, time_limit=120.0, exclusions=exclusions,
                       cover=bool(cover), pattern=pattern, rank=rk, anchors=an,
                       use_kernel=kernel is not None)
        want = oracle.gm_search_ref(S, K, ref_tabs, len(relaxed), rk, an,
                                    use_ex=exclusions and S.ex_start is not None, enum_all=enum_all,
                                    task_depth=td, cover_gate=4 if cover else 0, ascent=1,
                                    kernel=S.kernel if kernel is not None else None, pattern=pattern)
    finally:
        tab.free()
    assert {0: "exhausted", 1: "found"}[want["status"]] == got["status"]
    assert got["best"] == want["best"]
    if enum_all:
        assert sorted(got["solutions"]) == want["solutions"]
        assert got["nodes"] == want["nodes"], (got["nodes"], want["nodes"])
        st = want["stats"]
        assert (got["hits"], got["false_hits"], got["tasks"], got["cover_pruned"], got["ascent_pruned"]) == \
            (st[1], st[2], st[3], st[4], st[5])
        assert got["nsol_found"] == st[7]
    return got, want


@needs_ref
@pytest.mark.parametrize("labelfree", [False, True])
def test_codes_against_reference_search(labelfree):
    for p, d in _codes():
        S = search.Searcher(p, labelfree=labelfree)
        S.build_exclusions()
        for kt, cover in ((2, 0), (2, 1), (3, 1)):
            print("case search vs reference", p.name, p.ncols, "d", d, "kt", kt, "cover", cover, "labelfree",
                  labelfree, flush=True)
            for K in range(1, d + 1):
                got, _ = _compare(S, p, K, kt, labelfree, cover=cover)
                assert (got["best"] is None) == (K < d)
            assert got["best"] == d


@needs_ref
def test_first_solution_mode_against_reference():
    """enum_all = False: the run stops at its first solution; status and best weight agree."""
    for p, d in _codes()[:4]:
        S = search.Searcher(p)
        S.build_exclusions()
        print("case search first solution", p.name, flush=True)
        got, want = _compare(S, p, d, 2, False, enum_all=False)
        assert got["status"] == "found" and got["best"] == d and search.verify(p, got["solutions"][0])


@needs_ref
def test_kernel_sets_pattern_and_relaxed_tables_against_reference():
    p = _codes()[0][0]                       # code_bb72
    S = search.Searcher(p)
    S.build_exclusions()
    # kernel mode: zero-syndrome zero-logical sets up to weight 6, device vs reference
    sets, r = S.kernel_sets(6)
    want = oracle.gm_search_ref(S, 6, [], -1, S.rank, S.anchors, use_ex=False, enum_all=True, task_depth=2,
                                kernel_mode=True)
    assert sorted(s for s in want["solutions"] if len(s) >= 3) == sorted(sets)
    assert r["nodes"] == want["nodes"]
    print("case search kernel sets", len(sets), flush=True)
    S.set_kernel(sets)
    for K in range(1, 9):
        got, _ = _compare(S, p, K, 2, False, kernel=True)
        if got["best"] is not None:
            break
    # pattern budget: classes by column index mod 5, classes 0-2 on pattern, one extra column
    cls = np.array([j % 5 if j % 7 else -1 for j in range(p.ncols)], dtype=np.int32)
    onpat = np.array([1, 1, 1, 0, 0], dtype=np.uint8)
    print("case search pattern", flush=True)
    for K in range(1, 9):
        _compare(S, p, K, 2, False, pattern=(cls, onpat, 1))
    # a relaxed table: the problem with its last rows dropped, columns mapped by masking
    keep = (1 << (p.nrows - 3)) - 1
    img = [x & keep for x in p.h]
    rp = Problem("relaxed", p.nrows - 3, p.nobs, img, list(p.l), [()] * (p.nrows - 3))
    rtab = tables.Table(rp, 2)
    try:
        rzh, rzl = rtab.hashes_for(img)
        rlg = np.array(rp.l, dtype=np.uint64)
        print("case search relaxed table", flush=True)
        for K in range(1, 9):
            got, _ = _compare(S, p, K, 1, False, relaxed=[(rtab, rzh, rzl, rlg)])
            if got["best"] is not None:
                break
    finally:
        rtab.free()


@needs_ref
def test_anchors_and_ranks_against_reference():
    p = _codes()[0][0]
    rng = random.Random(4)
    rank = [rng.randrange(3) for _ in range(p.ncols)]
    anchors = sorted(range(p.ncols), key=lambda j: (rank[j], j))
    S = search.Searcher(p, rank=rank, anchors=anchors, labelfree=True)
    S.build_exclusions()
    print("case search anchors and ranks", flush=True)
    for K in range(1, 9):
        got, _ = _compare(S, p, K, 2, True, cover=1)
        if got["best"] is not None:
            break


def test_time_limit_reports_timeout():
    p = problem.quasi_cyclic(6, 6, seed=3001)
    S = search.Searcher(p)
    tab = tables.Table(p, 1)
    S.add_table(tab, exact=True)
    try:
        print("case search time limit", flush=True)
        r = S.search(9, enum_all=True, time_limit=1e-4)
        assert r["status"] == "timeout"
    finally:
        tab.free()
