"""Distance problems: columns as Python-int bitmasks, and the seeded synthetic instances the benchmarks use.

A problem is a list of columns; column j has a detector part h[j] (bitmask over rows), a logical part l[j] (bitmask
over observables) and a cost w[j] (1 unless given).  The distance is

    d = min { sum_j e_j w_j : XOR_j e_j h[j] = 0,  XOR_j e_j l[j] != 0 }

(the reference's model, cdist/problem.py:1-35, where every column costs one fault; integer costs are this
library's extension).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    name: str
    nrows: int
    nobs: int
    h: list
    l: list
    row_meta: list = field(default_factory=list)
    col_meta: list = field(default_factory=list)
    w: list = None                     # integer cost per column; None = unit costs

    @property
    def ncols(self):
        return len(self.h)

    def weight(self, j):
        return 1 if self.w is None else int(self.w[j])

    def check(self, support):
        """(syndrome, logical) of a set of columns (cdist/problem.py:30-38)."""
        s = lg = 0
        for j in support:
            s ^= self.h[j]
            lg ^= self.l[j]
        return s, lg

    def cost(self, support):
        return sum(self.weight(j) for j in support)

    def words(self):
        """The columns as the planner's uint64 bitsets (trellis.column_words), computed once per column list."""
        from .trellis import column_words
        c = self.__dict__.get("_words")
        if c is None or c[0] is not self.h:
            c = self.__dict__["_words"] = (self.h, column_words(self.h, self.nrows))
        return c[1]


def dedupe(h, l, w=None):
    """Drop all-zero columns and merge identical (h, l) columns, keeping the least cost (cdist/problem.py:47-64 for
    unit costs).  Returns (h, l, w) with w None when no costs were given."""
    seen = {}
    oh, ol, ow = [], [], []
    for j, (a, b) in enumerate(zip(h, l)):
        if a == 0 and b == 0:
            continue
        c = 1 if w is None else int(w[j])
        k = seen.get((a, b))
        if k is not None:
            ow[k] = min(ow[k], c)
            continue
        seen[(a, b)] = len(oh)
        oh.append(a)
        ol.append(b)
        ow.append(c)
    return oh, ol, (None if w is None else ow)


def verify_witness(p, cols, value, obs=None):
    """True when the columns form a solution of cost `value`: H e = 0, L e != 0 (or bit obs of L e set) and no
    column repeated (search.verify, cdist/search.py:381-384; common.check_witness, bench/common.py:138-147)."""
    if cols is None or len(set(cols)) != len(cols):
        return False
    s, lg = p.check(cols)
    ok_l = ((lg >> obs) & 1) if obs is not None else lg != 0
    return s == 0 and bool(ok_l) and p.cost(cols) == value


# ---------------------------------------------------------------------------------------------- generators
def banded(R, N, kmin, kmax, W, nobs, rho, weights=None, seed=0, name=None):
    """Banded random detector error model.  Column j flips k_j ~ U{kmin..kmax} distinct detectors drawn uniformly
    from the window [s_j, s_j + W), s_j = round(j (R - W) / (N - 1)); each of the nobs logical bits is set with
    probability rho; weights: None (unit) or (lo, hi) for costs ~ U{lo..hi}.  Zero and duplicate columns are merged
    (least cost kept).  The natural column order has cut rank <= W."""
    rng = np.random.default_rng(seed)
    starts = np.rint(np.arange(N) * (R - W) / max(1, N - 1)).astype(np.int64)
    k = rng.integers(kmin, kmax + 1, size=N)
    picks = np.argsort(rng.random((N, W)), axis=1)           # a random permutation of the window per column
    chosen = np.zeros((N, W), dtype=bool)
    np.put_along_axis(chosen, picks, np.arange(W)[None, :] < np.minimum(k, W)[:, None], axis=1)
    lbits = rng.random((N, nobs)) < rho
    wts = rng.integers(weights[0], weights[1] + 1, size=N) if weights is not None else np.ones(N, dtype=np.int64)
    packed = np.packbits(chosen, axis=1, bitorder="little")
    h = [int.from_bytes(packed[j].tobytes(), "little") << int(starts[j]) for j in range(N)]
    l = (lbits.astype(np.int64) << np.arange(nobs)).sum(axis=1).tolist() if nobs else [0] * N
    hh, ll, ww = dedupe(h, l, wts.tolist())
    if weights is None:
        ww = None
    return Problem(name or "banded_R%d_N%d_W%d_s%d" % (R, N, W, seed), R, nobs, hh, ll, w=ww)


def _rref(rows):
    """Reduced row echelon basis (list of ints) and pivot bit per row."""
    basis, pivots = [], []
    for v in rows:
        for b, p in zip(basis, pivots):
            if (v >> p) & 1:
                v ^= b
        if v:
            p = v.bit_length() - 1
            for i in range(len(basis)):
                if (basis[i] >> p) & 1:
                    basis[i] ^= v
            basis.append(v)
            pivots.append(p)
    return basis, pivots


def _nullspace(rows, n):
    """Basis of {x in GF(2)^n : <r, x> = 0 for every row r}."""
    basis, pivots = _rref(rows)
    pivset = set(pivots)
    out = []
    for f in range(n):
        if f in pivset:
            continue
        x = 1 << f
        for b, p in zip(basis, pivots):
            if (b >> f) & 1:
                x |= 1 << p
        out.append(x)
    return out


def quotient_basis(kernel_of, modulo, n):
    """Basis of ker(kernel_of) / rowspace(modulo) (the logical operators of a CSS code, as in
    papers/stabcode.py:72-93)."""
    basis, pivots = _rref(list(modulo))
    out = []
    for v in _nullspace(kernel_of, n):
        for b, p in zip(basis, pivots):
            if (v >> p) & 1:
                v ^= b
        if v:
            p = v.bit_length() - 1
            for i in range(len(basis)):
                if (basis[i] >> p) & 1:
                    basis[i] ^= v
            basis.append(v)
            pivots.append(p)
            out.append(v)
    return out


def two_block_matrices(l, m, A, B):
    """H_X = [A | B], H_Z = [B^T | A^T] on the l x m torus, A and B as lists of exponent pairs (a, b) of monomials
    x^a y^b (the construction of papers/stabcode.py:237-258)."""
    N = l * m

    def circ(P, transpose):
        out = []
        for i in range(l):
            for j in range(m):
                v = 0
                for a, b in P:
                    ii, jj = ((i - a) % l, (j - b) % m) if transpose else ((i + a) % l, (j + b) % m)
                    v ^= 1 << (ii * m + jj)
                out.append(v)
        return out
    a, b, at, bt = circ(A, False), circ(B, False), circ(A, True), circ(B, True)
    hx = [a[r] | (b[r] << N) for r in range(N)]
    hz = [bt[r] | (at[r] << N) for r in range(N)]
    return hx, hz, 2 * N


def css_problem(hx, hz, n, name="css"):
    """Basis-Z code-capacity problem: columns of H_Z as detectors, logical rows spanning ker(H_X) / rowspace(H_Z)
    (papers/stabcode.py:117-125)."""
    lz = quotient_basis(hx, hz, n)
    h = [sum(1 << i for i, r in enumerate(hz) if (r >> j) & 1) for j in range(n)]
    lg = [sum(1 << i for i, r in enumerate(lz) if (r >> j) & 1) for j in range(n)]
    hh, ll, _ = dedupe(h, lg)
    return Problem(name, len(hz), len(lz), hh, ll)


def quasi_cyclic(l, m, seed=0, terms=3, name=None):
    """Seeded two-block quasi-cyclic code: A and B are sums of `terms` distinct random monomials x^a y^b.  Draws
    again (same generator) until the code has at least one logical qubit."""
    rng = np.random.default_rng(seed)
    for _ in range(1000):
        cells = [(a, b) for a in range(l) for b in range(m)]
        A = [cells[i] for i in rng.choice(len(cells), size=terms, replace=False)]
        B = [cells[i] for i in rng.choice(len(cells), size=terms, replace=False)]
        hx, hz, n = two_block_matrices(l, m, A, B)
        p = css_problem(hx, hz, n, name or "qc_l%d_m%d_s%d" % (l, m, seed))
        if p.nobs > 0:
            p.row_meta = [("A", A), ("B", B)]
            return p
    raise ValueError("no two-block code with logical qubits for this seed")


def chain(n, weights=None):
    """The chain check of the 8-bit boundary tests: e_0 = ... = e_{n-1} with logical e_0, distance n (unit) or
    sum(weights) (tests/test_false_bounds.py:171-174)."""
    h = [1] + [(1 << (j - 1)) | (1 << j) for j in range(1, n - 1)] + [1 << (n - 2)]
    return Problem("chain%d" % n, n - 1, 1, h, [1] + [0] * (n - 1), w=weights)
