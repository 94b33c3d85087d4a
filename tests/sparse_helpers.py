"""Random small problems of the reference's sparse tests (tests/test_sparse_dp.py:20-46)."""
import itertools

from tropidist_amd.problem import Problem


def brute(p, obs):
    for w in range(1, p.ncols + 1):
        for cols in itertools.combinations(range(p.ncols), w):
            s = lg = 0
            for j in cols:
                s ^= p.h[j]
                lg ^= p.l[j]
            if s == 0 and (lg >> obs) & 1:
                return w
    return None


def check_witness(p, cols, value, obs=None):
    s = lg = 0
    for j in cols:
        s ^= p.h[j]
        lg ^= p.l[j]
    assert s == 0 and len(cols) == value == len(set(cols))
    assert ((lg >> obs) & 1) if obs is not None else lg != 0


def random_problem(rng):
    n, m, k = rng.randint(2, 12), rng.randint(1, 7), rng.randint(1, 2)
    dens = rng.choice([0.2, 0.35, 0.5])
    h = [sum(1 << r for r in range(m) if rng.random() < dens) for _ in range(n)]
    l = [sum(1 << r for r in range(k) if rng.random() < 0.4) for _ in range(n)]
    return Problem("rand", m, k, h, l)
