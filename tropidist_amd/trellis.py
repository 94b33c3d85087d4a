"""Linear rank decomposition (syndrome trellis) of a distance query, and its DP plan.

Restates the planner of the reference (cdist/trellis.py:16-193): for a column order, rows of the check matrix are
brought into minimal span form (distinct starts, distinct ends), the target-1 row is shortened against them, and
the plan walks the columns, opening a table bit for each row that starts, applying one min-plus butterfly per
column over the bits of the rows it touches, and closing each row that ends by keeping the half of the states whose
bit equals the row's target.

Extension over the reference: every column may carry an integer weight (>= 1).  A unit column emits
("update", mask) exactly as the reference does; a column of weight w != 1 emits ("update", mask, w).  Plan.work
and Plan.peak_bits do not depend on the weights.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import itertools

import numpy as np

OP_EXPAND, OP_UPDATE, OP_SELECT = 0, 1, 2


def lowbit(x):
    """Index of the lowest set bit (cdist/trellis.py:16-17); -1 for 0."""
    return (x & -x).bit_length() - 1


def rank_profile(cols, order):
    """r[i] = GF(2) rank of the first i columns of the order (cdist/trellis.py:20-36)."""
    pivots = {}
    out = [0]
    for j in order:
        v = cols[j]
        while v:
            top = v.bit_length() - 1
            p = pivots.get(top)
            if p is None:
                pivots[top] = v
                break
            v ^= p
        out.append(len(pivots))
    return out


def cut_profile(cols, order):
    """Cut ranks lambda_i = rank(past) + rank(future) - rank(all), i = 0..N (cdist/trellis.py:39-44)."""
    fwd = rank_profile(cols, order)
    bwd = rank_profile(cols, list(reversed(order)))[::-1]
    return [a + b - fwd[-1] for a, b in zip(fwd, bwd)]


def column_words(cols, nrows):
    """Columns as a (ncols, ceil(nrows / 64)) uint64 array (the bitset layout of the C++ planner)."""
    rw = max(1, (nrows + 63) // 64)
    if not cols:
        return np.zeros((1, rw), dtype=np.uint64)
    nb = 8 * rw
    try:                                      # map calls int.to_bytes without a Python frame per column
        raw = b"".join(map(int.to_bytes, cols, itertools.repeat(nb), itertools.repeat("little")))
    except TypeError:                         # columns that are not Python ints (numpy scalars)
        raw = b"".join(int(c).to_bytes(nb, "little") for c in cols)
    return np.frombuffer(raw, dtype=np.uint64).reshape(len(cols), rw).copy()


def parity_targets(l, obs_masks):
    """targets[q, j] = parity of l[j] & obs_masks[q] (uint8), for the C++ planner."""
    nq = len(obs_masks)
    out = np.zeros((max(1, nq), max(1, len(l))), dtype=np.uint8)
    if not l:
        return out
    if max(l) < (1 << 64) and max(obs_masks, default=0) < (1 << 64):
        la = np.asarray(l, dtype=np.uint64)
        for q, m in enumerate(obs_masks):
            out[q, :len(l)] = (np.bitwise_count(la & np.uint64(m)) & 1).astype(np.uint8)
        return out
    for q, m in enumerate(obs_masks):
        for j, x in enumerate(l):
            out[q, j] = bin(x & m).count("1") & 1
    return out


def cut_profile_native(cols, order, nrows, words=None):
    """cut_profile by libtdp's C++ rank profile (tdp_rank_profile); words: column_words(cols, nrows) if known."""
    import ctypes
    from . import _lib
    L = _lib.lib()
    hw = column_words(cols, nrows) if words is None else words
    n = len(order)
    fwd_o = np.ascontiguousarray(order, dtype=np.int32)
    bwd_o = np.ascontiguousarray(fwd_o[::-1])
    fwd = np.zeros(n + 1, dtype=np.int32)
    bwd = np.zeros(n + 1, dtype=np.int32)
    for o, r in ((fwd_o, fwd), (bwd_o, bwd)):
        _lib.check(L.tdp_rank_profile(hw.ctypes.data_as(ctypes.c_void_p), len(cols), hw.shape[1],
                                      o.ctypes.data_as(ctypes.c_void_p), n, r.ctypes.data_as(ctypes.c_void_p)),
                   "tdp_rank_profile")
    return (fwd + bwd[::-1] - fwd[-1]).tolist()


def rows_from_columns(cols, order, nrows):
    """Row r as a bitmask over column positions of the order (cdist/trellis.py:47-57)."""
    rows = [0] * nrows
    for pos, j in enumerate(order):
        v = cols[j]
        while v:
            low = v & -v
            rows[low.bit_length() - 1] |= 1 << pos
            v ^= low
    return rows


def minimal_span_form(rows):
    """Row reduction to distinct starts and distinct ends; dependent rows vanish (cdist/trellis.py:60-97).

    Pass 1 inserts rows keyed by their start; on a clash the row that ends earlier keeps the slot and the sum moves
    on.  Pass 2 walks rows by decreasing start and keys them by their end; on a clash the row starting earlier
    absorbs the other (its start stays, its end moves left)."""
    starts = {}
    for r in rows:
        while r:
            s = lowbit(r)
            held = starts.get(s)
            if held is None:
                starts[s] = r
                break
            if r.bit_length() < held.bit_length():
                starts[s], r = r, held
            r ^= starts[s]
    ends = {}
    for r in sorted(starts.values(), key=lowbit, reverse=True):
        while True:
            e = r.bit_length()
            held = ends.get(e)
            if held is None:
                ends[e] = r
                break
            if lowbit(held) > lowbit(r):
                r ^= held
            else:
                ends[e], r = r, held ^ r
    out = list(ends.values())
    assert len({lowbit(r) for r in out}) == len(out) and len({r.bit_length() for r in out}) == len(out)
    return out


def reduce_target_row(lrow, hrows):
    """Shorten the target-1 row by constraint rows: clear clashing starts, then clashing ends of rows that start no
    earlier, re-clearing starts after each step (cdist/trellis.py:100-122).  0 means the query is infeasible."""
    by_start = {lowbit(r): r for r in hrows}
    by_end = {r.bit_length(): r for r in hrows}

    def clear_start(x):
        while x:
            q = by_start.get(lowbit(x))
            if q is None:
                break
            x ^= q
        return x

    lrow = clear_start(lrow)
    while lrow:
        q = by_end.get(lrow.bit_length())
        if q is None or lowbit(q) < lowbit(lrow):
            break
        lrow = clear_start(lrow ^ q)
    return lrow


@dataclass
class Plan:
    """Op stream of the dense DP: ("expand",) | ("update", mask[, weight]) | ("select", pos, target)
    (cdist/trellis.py:125-144)."""
    ops: list
    peak_bits: int
    widths: list
    ncols: int
    feasible: bool = True
    columns: list = field(default_factory=list)     # column position (in the order) of each update op
    _encoded: object = field(default=None, init=False, repr=False, compare=False)
    _encoded_key: object = field(default=None, init=False, repr=False, compare=False)

    def work(self):
        """Table entries touched by updates: sum over updates of the table size at that update."""
        tot, bits = 0, 0
        for op in self.ops:
            if op[0] == "expand":
                bits += 1
            elif op[0] == "select":
                bits -= 1
            else:
                tot += 1 << bits
        return tot

    def encode(self):
        """The op stream as an (nops, 3) uint64 array of (code, a, b): expand (0, 0, 0), update (1, mask, weight),
        select (2, pos, target) -- the format of the C-ABI plan executors.  The array is read-only and kept on the
        plan while `ops` is the same list of the same length, so a plan run many times is encoded once (assign a
        new list to change a plan's ops; edits in place are not seen)."""
        key = self._encoded_key
        if key is not None and key[0] is self.ops and key[1] == len(self.ops):
            return self._encoded
        flat = []
        put = flat.extend
        for op in self.ops:
            kind = op[0]
            if kind == "expand":
                put((OP_EXPAND, 0, 0))
            elif kind == "update":
                put((OP_UPDATE, op[1], op[2] if len(op) > 2 else 1))
            else:
                put((OP_SELECT, op[1], op[2]))
        out = np.fromiter(flat, dtype=np.uint64, count=len(flat)).reshape(-1, 3)
        out.setflags(write=False)
        self._encoded, self._encoded_key = out, (self.ops, len(self.ops))   # holds the list: its id stays its own
        return out

    def max_weight(self):
        return max((op[2] if len(op) > 2 else 1 for op in self.ops if op[0] == "update"), default=1)

    @property
    def nops(self):
        return len(self.ops)


def as_plan(plan):
    """A Plan for any object with the reference's plan attributes (ops, peak_bits, feasible; cdist/trellis.py:125-144),
    so the executors accept plans made by the reference's own planner; Plans pass through unchanged."""
    if isinstance(plan, Plan):
        return plan
    ops = [tuple(op) for op in plan.ops]
    return Plan(ops, int(plan.peak_bits), list(getattr(plan, "widths", []) or []),
                int(getattr(plan, "ncols", 0) or 0), bool(plan.feasible))


def make_plan(hrows, lrow, ncols, weights=None):
    """DP schedule for constraint rows (target 0) and one target-1 row (cdist/trellis.py:147-182); weights[j] is the
    cost of the column at position j (default 1)."""
    if lrow == 0:
        return Plan([], 0, [], ncols, feasible=False)
    rows = list(hrows) + [lrow]
    target = [0] * len(hrows) + [1]
    opening, closing = {}, {}
    for idx, r in enumerate(rows):
        opening.setdefault(lowbit(r), []).append(idx)
        closing.setdefault(r.bit_length() - 1, []).append(idx)
    active, ops, widths, columns, peak = [], [], [], [], 0
    for j in range(ncols):
        for idx in opening.get(j, ()):
            active.append(idx)
            ops.append(("expand",))
        peak = max(peak, len(active))
        mask = 0
        for pos, idx in enumerate(active):
            if (rows[idx] >> j) & 1:
                mask |= 1 << pos
        if mask:                       # a column touching no active row is a free zero-syndrome fault
            w = 1 if weights is None else int(weights[j])
            ops.append(("update", mask) if w == 1 else ("update", mask, w))
            columns.append(j)
        for pos in sorted((active.index(idx) for idx in closing.get(j, ())), reverse=True):
            ops.append(("select", pos, target[active[pos]]))
            active.pop(pos)
        widths.append(len(active))
    assert not active
    return Plan(ops, peak, widths, ncols, columns=columns)


def target_row(l, obs_mask, order):
    """Bitmask over positions of the columns whose logical label has odd parity with obs_mask."""
    lrow = 0
    for pos, j in enumerate(order):
        if bin(l[j] & obs_mask).count("1") & 1:
            lrow |= 1 << pos
    return lrow


def build_plan(h, l, obs_mask, order, nrows, weights=None):
    """Plan for H e = 0, (obs_mask . L) e = 1 in the given order (cdist/trellis.py:185-193); weights are per
    column of the problem (not per position)."""
    hrows = minimal_span_form(rows_from_columns(h, order, nrows))
    lrow = reduce_target_row(target_row(l, obs_mask, order), hrows)
    pw = None if weights is None else [weights[j] for j in order]
    return make_plan(hrows, lrow, len(order), pw)


class NativePlan(Plan):
    """A Plan made by the C++ planner: the op stream is kept encoded and decoded only when .ops is read."""

    def __init__(self, encoded, peak_bits, ncols, work, columns, feasible=True):
        self._encoded = encoded
        self._work = int(work)
        self._decoded = None
        Plan.__init__(self, None, peak_bits, [], ncols, feasible, list(columns))

    @property
    def nops(self):
        return len(self._encoded)

    @property
    def ops(self):
        if self._decoded is None:
            out = []
            for code, a, b in self._encoded.tolist():
                if code == OP_EXPAND:
                    out.append(("expand",))
                elif code == OP_UPDATE:
                    out.append(("update", a) if b == 1 else ("update", a, b))
                else:
                    out.append(("select", a, b))
            self._decoded = out
        return self._decoded

    @ops.setter
    def ops(self, value):
        pass

    def encode(self):
        return self._encoded

    def work(self):
        return self._work

    def max_weight(self):
        return int(self._encoded[self._encoded[:, 0] == OP_UPDATE, 2].max(initial=1))


def _words(x, nw):
    return np.frombuffer(int(x).to_bytes(8 * nw, "little"), dtype=np.uint64)


def build_plans_native(h, l, obs_masks, order, nrows, weights=None, words=None):
    """Plans of several queries (one per obs_mask) by the C++ planner of libtdp (tdp_plan_build); identical op
    streams to make_plan.  Plans with more than 63 active rows are refused (ValueError).  words: column_words(h,
    nrows) when the caller has it."""
    import ctypes
    from . import _lib
    L = _lib.lib()
    ncols, nq = len(order), len(obs_masks)
    hw = column_words(h, nrows) if words is None else words
    targets = parity_targets(l, obs_masks)
    order_a = np.asarray(order, dtype=np.int32)
    w = None if weights is None else np.asarray(weights, dtype=np.int32)
    cap = ncols + 2 * (nrows + 1) + 4
    # the planner writes every field of every query and the first nops[q] ops / columns it reports
    ops = np.empty((max(1, nq), cap, 3), dtype=np.uint64)
    cols = np.empty((max(1, nq), cap), dtype=np.int32)
    nops = np.zeros(max(1, nq), dtype=np.int64)
    peak = np.zeros(max(1, nq), dtype=np.int32)
    work = np.zeros(max(1, nq), dtype=np.float64)
    feas = np.zeros(max(1, nq), dtype=np.int32)
    p_ = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = L.tdp_plan_build(p_(hw), ncols, nrows, p_(order_a), p_(targets), nq, None if w is None else p_(w),
                          p_(ops), cap, p_(nops), p_(peak), p_(work), p_(cols), p_(feas))
    if rc == _lib.TDP_EINVAL:
        raise ValueError(L.tdp_last_error().decode())
    _lib.check(rc, "tdp_plan_build")
    out = []
    for q in range(nq):
        if not feas[q]:
            out.append(Plan([], 0, [], ncols, feasible=False))
            continue
        n = int(nops[q])
        nupd = int((ops[q, :n, 0] == OP_UPDATE).sum())
        out.append(NativePlan(ops[q, :n].copy(), int(peak[q]), ncols, work[q], cols[q, :nupd].tolist()))
    return out
