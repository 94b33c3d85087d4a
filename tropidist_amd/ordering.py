"""Column orders with a small trellis width, for callers without one of their own.

greedy_straddle_order grows the processed set one column at a time, taking the column that opens the fewest new
rows net of the rows it closes (a row is open while it has processed and unprocessed columns); ties go to the
column with the most already-open rows, then to the lowest index.  The number of open rows bounds the cut rank, so
the order keeps the dense table and the sparse trellis narrow.  This is a host heuristic of this library (the
reference's own heuristics, cdist/ordering.py, optimise the GF(2) cut rank directly); any order is exact, only the
cost depends on it."""
from __future__ import annotations

from . import trellis


def greedy_straddle_order(cols, nrows, start=0):
    n = len(cols)
    if n == 0:
        return []
    row_cols = [[] for _ in range(nrows)]
    for j, c in enumerate(cols):
        x = c
        while x:
            low = x & -x
            row_cols[low.bit_length() - 1].append(j)
            x ^= low
    remaining_in_row = [len(rc) for rc in row_cols]
    opened = [False] * nrows
    done = [False] * n
    order = []
    # candidate scores only change for columns sharing a row with the chosen one
    def score(j):
        new = closes = touched = 0
        x = cols[j]
        while x:
            low = x & -x
            r = low.bit_length() - 1
            x ^= low
            if not opened[r]:
                new += remaining_in_row[r] > 1
            else:
                touched += 1
            if remaining_in_row[r] == 1 and opened[r]:
                closes += 1
        return (new - closes, -touched, j)

    frontier = set()
    current = start if 0 <= start < n else 0
    while True:
        order.append(current)
        done[current] = True
        x = cols[current]
        while x:
            low = x & -x
            r = low.bit_length() - 1
            x ^= low
            opened[r] = True
            remaining_in_row[r] -= 1
            for j in row_cols[r]:
                if not done[j]:
                    frontier.add(j)
        frontier.discard(current)
        if len(order) == n:
            break
        if frontier:
            current = min(frontier, key=score)
        else:
            current = next(j for j in range(n) if not done[j])
    return order


def greedy_order_native(cols, nrows, words=None, start=0):
    """greedy_straddle_order by libtdp's C++ planner (tdp_greedy_order): the same order, without the Python loop.
    words: trellis.column_words(cols, nrows) when the caller has them."""
    import ctypes

    import numpy as np

    from . import _lib
    n = len(cols)
    if n == 0:
        return []
    hw = words if words is not None else trellis.column_words(cols, nrows)
    out = np.zeros(n, dtype=np.int32)
    _lib.check(_lib.lib().tdp_greedy_order(hw.ctypes.data_as(ctypes.c_void_p), n, nrows, start,
                                           out.ctypes.data_as(ctypes.c_void_p)), "tdp_greedy_order")
    return out.tolist()


def order_width(cols, order):
    """Largest cut rank of an order."""
    return max(trellis.cut_profile(cols, order))
