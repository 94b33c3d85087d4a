"""Shared generators for the tests (random small problems in the patterns of the reference's tests)."""
import gzip
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def random_problem(rng):
    """tests/test_dp.py:26-33 of the reference: m <= 9 rows, n <= 14 columns, k <= 3 observables."""
    m = rng.randint(1, 9)
    n = rng.randint(1, 14)
    k = rng.randint(1, 3)
    dens = rng.choice([0.15, 0.3, 0.5])
    h = [sum(1 << i for i in range(m) if rng.random() < dens) for _ in range(n)]
    l = [sum(1 << i for i in range(k) if rng.random() < 0.3) for _ in range(n)]
    return m, k, h, l


def dense_fixtures():
    with gzip.open(os.path.join(HERE, "golden", "dense_audit.json.gz"), "rt") as f:
        out = json.load(f)
    for r in out:
        r["h"] = [int(x, 16) for x in r["h"]]
        r["l"] = [int(x, 16) for x in r["l"]]
    return out
