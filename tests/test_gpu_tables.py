"""GPU parity of the MITM lookup tables: device build and batched probes against the oracle's table."""
import random

import numpy as np
import pytest

import oracle
from test_oracle_tables import random_problem
from tropidist_amd import problem, tables

pytestmark = pytest.mark.gpu


def _check_against_oracle(p, k, seed, rng, nq=64):
    t = tables.Table(p, k, seed=seed)
    lg = np.array(p.l, dtype=np.uint64)
    h, count, free = oracle.tab_build(t.zh, t.zl, lg, k, t.log2slots, int(t.wide))
    assert t.entries == count
    qs = []
    for _ in range(nq):
        sup = rng.sample(range(p.ncols), rng.randint(0, min(k + 1, p.ncols)))
        a = b = g = 0
        for j in sup:
            a ^= int(t.zh[j])
            b ^= int(t.zl[j])
            g ^= int(lg[j])
        qs.append((a, b, g ^ rng.choice([0, 1, 2]), rng.randint(0, k)))
    got = t.lookup([q[0] for q in qs], [q[1] for q in qs], [q[2] for q in qs], [q[3] for q in qs])
    want = [oracle.tab_lookup(h, *q) for q in qs]
    assert list(got) == want
    free(h)
    t.free()


def test_random_tables_equal_oracle():
    print("case tables random", flush=True)
    rng = random.Random(31)
    for trial in range(60):
        p = random_problem(rng, nobs=rng.choice([2, 20]))
        _check_against_oracle(p, rng.randint(1, 3), trial, rng)


def test_forced_collision_on_gpu():
    print("case tables forced collision", flush=True)
    from unittest.mock import patch
    p = problem.Problem("collide", 1, 2, [1, 1], [1, 2])
    with patch.object(tables, "row_hashes", lambda nrows, seed: ([0] * nrows, [0] * nrows)):
        for labelfree in (False,):
            t = tables.Table(p, 2, labelfree=labelfree)
            assert t.entries == 4
            assert list(t.lookup([0, 0], [0, 0], [0, 1], [2, 0])) == [1, 0]
            t.free()


@pytest.mark.parametrize("k", [2, 3])
def test_quasi_cyclic_288_columns(k):
    """n = 288 columns: k = 2 inserts 41,617 subsets, k = 3 inserts 3,981,553; counts and probes equal the oracle."""
    print("case tables qc288 k", k, flush=True)
    p = problem.quasi_cyclic(12, 12, seed=1)
    _check_against_oracle(p, k, 12345, random.Random(k), nq=256)
