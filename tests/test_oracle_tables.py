"""MITM lookup tables: the oracle's table equals the reference's compiled gmitm.c in size, its lookups equal brute
force, and the row hashes are the reference's (Python's Mersenne Twister, cdist/search.py:83-85)."""
import random

import numpy as np
import pytest

import oracle
from tropidist_amd import tables
from tropidist_amd.problem import Problem

needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")


def random_problem(rng, nobs=None):
    m, n = rng.randint(2, 10), rng.randint(2, 14)
    k = nobs or rng.randint(1, 3)
    h = [sum(1 << i for i in range(m) if rng.random() < rng.choice([0.2, 0.35])) for _ in range(n)]
    l = [sum(1 << i for i in range(k) if rng.random() < 0.3) for _ in range(n)]
    return Problem("rand", m, k, h, l)


def test_row_hashes_are_the_reference_stream():
    rh = tables.row_hashes(3, 12345)
    rng = random.Random(12345)
    assert rh == ([rng.getrandbits(64) for _ in range(3)], [rng.getrandbits(64) for _ in range(3)])
    zh, zl = tables.column_hashes([0b101, 0b011], rh)
    assert int(zh[0]) == rh[0][0] ^ rh[0][2] and int(zl[1]) == rh[1][0] ^ rh[1][1]


@needs_ref
def test_counts_equal_reference_c():
    rng = random.Random(17)
    for t in range(200):
        p = random_problem(rng, nobs=rng.choice([2, 20]))
        k = rng.randint(1, 3)
        zh, zl = tables.column_hashes(p.h, tables.row_hashes(p.nrows, t))
        lg = np.array(p.l, dtype=np.uint64)
        wide = int(p.nobs > 16)
        log2 = tables.table_log2slots(p.ncols, k)
        a, ca, fa = oracle.tab_build(zh, zl, lg, k, log2, wide, "oracle")
        b, cb, fb = oracle.tab_build(zh, zl, lg, k, log2, wide, "ref")
        assert ca == cb, t
        fa(a)
        fb(b)


def test_lookups_equal_brute_force():
    rng = random.Random(23)
    for t in range(80):
        p = random_problem(rng, nobs=rng.choice([2, 20]))
        k = rng.randint(1, 3)
        zh, zl = tables.column_hashes(p.h, tables.row_hashes(p.nrows, t))
        lg = np.array(p.l, dtype=np.uint64)
        wide = int(p.nobs > 16)
        h, _, free = oracle.tab_build(zh, zl, lg, k, tables.table_log2slots(p.ncols, k), wide)
        for q in range(12):
            sup = rng.sample(range(p.ncols), rng.randint(0, min(4, p.ncols)))
            qzh = qzl = qlg = 0
            for j in sup:
                qzh ^= int(zh[j])
                qzl ^= int(zl[j])
                qlg ^= int(lg[j])
            qlg ^= rng.choice([0, 0, 1, 3])
            R = rng.randint(0, k)
            assert oracle.tab_lookup(h, qzh, qzl, qlg, R) == \
                oracle.brute_tab_lookup(zh, zl, lg, k, qzh, qzl, qlg, R, wide), (t, q)
        free(h)


def test_forced_collision_table():
    """tests/test_collision.py: every row hash zero, M = (1 1), labels 1 and 2: one fingerprint for all subsets;
    the lookup of the empty syndrome with label 0 finds cost 0 is excluded (same label) but cost 1 is present."""
    zh = np.zeros(2, np.uint64)
    zl = np.zeros(2, np.uint64)
    lg = np.array([1, 2], np.uint64)
    for wide in (0, 1):
        h, count, free = oracle.tab_build(zh, zl, lg, 2, 10, wide)
        assert count == 4                         # labels 0, 1, 2, 3
        assert oracle.tab_lookup(h, 0, 0, 0, 2) == 1 and oracle.tab_lookup(h, 0, 0, 1, 0) == 0
        free(h)
