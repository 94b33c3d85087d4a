"""Seeded synthetic workloads of the benchmarks (instance generators only; seed = 1000 * config + index).

  single_cpu()        48 detectors, 300 columns (1-4 detectors in a 10-row window), 1 logical row of density 0.1,
                      unit costs: the CPU reference case (width <= 14)
  single_large()      600 detectors, 6000 columns (2-6 detectors in a 25-row window), 4 logical rows of density
                      0.05, costs 1-3: 2^26-entry tables
  sparse_qc()         two-block quasi-cyclic code, l = m = 12 (288 columns)
  bivariate_bicycle_288()  the [[288,12,18]] bivariate bicycle code (l = m = 12, A = x^3 + y^2 + y^7,
                      B = y^3 + x + x^2), Z-basis code-capacity problem, with its two translation orbits as ranks
                      and anchors
  batch_scan(n)       n banded models, 200-1200 detectors, 5-15 faults per detector, width 10-22, half unit and half
                      costs 1-3
  serving_mix(n)      the request mix of the serving loop (70% dense width 10-20, 20% width 21-25, 10% sparse QC)
"""
from __future__ import annotations

import numpy as np

from . import problem


def single_cpu(seed=1000):
    return problem.banded(R=48, N=300, kmin=1, kmax=4, W=10, nobs=1, rho=0.1, seed=seed, name="cfg1_banded")


def single_large(seed=2000, W=25):
    return problem.banded(R=600, N=6000, kmin=2, kmax=6, W=W, nobs=4, rho=0.05, weights=(1, 3), seed=seed,
                          name="cfg2_banded")


def sparse_qc(seed=3000, l=12, m=12):
    return problem.quasi_cyclic(l, m, seed=seed, name="cfg3_qc")


def bivariate_bicycle_288():
    """(problem, rank, anchors): columns 0-143 form the first block and 144-287 the second; the l x m translations
    act transitively on each, so one anchor per block with rank = block index covers every solution."""
    hx, hz, n = problem.two_block_matrices(12, 12, [(3, 0), (0, 2), (0, 7)], [(0, 3), (1, 0), (2, 0)])
    p = problem.css_problem(hx, hz, n, "bb288")
    if p.ncols != 288:
        raise ValueError("bb288: expected 288 distinct columns, got %d" % p.ncols)
    return p, [j // 144 for j in range(288)], [0, 144]


def banded_instance(rng, seed, width, weighted, nobs=2):
    R = int(rng.integers(200, 1201))
    f = int(rng.integers(5, 16))
    N = max(width + 2, int(round(R * f / 4)))
    return problem.banded(R=R, N=N, kmin=2, kmax=6, W=width, nobs=nobs, rho=0.05,
                          weights=(1, 3) if weighted else None, seed=seed, name="scan_%d" % seed)


def batch_scan(n=512, wmin=10, wmax=22, seed=4000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w = int(rng.integers(wmin, wmax + 1))
        out.append(banded_instance(rng, seed + i + 1, w, weighted=bool(i % 2)))
    return out


def serving_mix(n=10000, seed=5000, dense_lo=(10, 20), dense_hi=(21, 25), qc=(6, 15), sparse_frac=0.10,
                wide_frac=0.20, K=(4, 10)):
    """Request descriptors (kind, parameters) -- problems are built lazily by the serving loop."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        u = rng.random()
        s = seed + i + 1
        if u < sparse_frac:
            # torus sizes with a side divisible by 3: the two-block codes of the other sizes rarely encode qubits
            while True:
                l, m = int(rng.integers(qc[0], qc[1] + 1)), int(rng.integers(qc[0], qc[1] + 1))
                if l % 3 == 0 or m % 3 == 0:
                    break
            reqs.append(("sparse", {"l": l, "m": m, "seed": s, "K": int(rng.integers(K[0], K[1] + 1))}))
        elif u < sparse_frac + wide_frac:
            reqs.append(("dense", {"width": int(rng.integers(dense_hi[0], dense_hi[1] + 1)), "seed": s,
                                   "weighted": bool(rng.random() < 0.5)}))
        else:
            reqs.append(("dense", {"width": int(rng.integers(dense_lo[0], dense_lo[1] + 1)), "seed": s,
                                   "weighted": bool(rng.random() < 0.5)}))
    return reqs
