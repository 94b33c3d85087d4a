"""Meet-in-the-middle lookup tables on the GPU: every column subset of size <= k keyed by syndrome hash (the
tables of the reference's cdist/search.py:83-142 and cdist/gmitm.c:36-186).

    row_hashes(nrows, seed), column_hashes(h, rh)   64-bit XOR-linear row hashes (a set's hash is the XOR of its
                                                    columns'), drawn from random.Random(seed) as the reference does
    Table(p, k, seed=12345, load=0.5, labelfree=False)
        .lookup(zh, zl, lg, R)   batched: least cost <= R of a stored subset with the same syndrome fingerprint
                                 and a different logical label, -1 if none
        .entries, .log2slots, .build_seconds, .free()

A fingerprint collision can only add a spurious hit, never hide a solution; callers re-derive hits."""
from __future__ import annotations

import math
import random
import time

import numpy as np

from . import _lib


def row_hashes(nrows, seed):
    """Two 64-bit hashes per row (cdist/search.py:83-85)."""
    rng = random.Random(seed)
    return ([rng.getrandbits(64) for _ in range(nrows)], [rng.getrandbits(64) for _ in range(nrows)])


def column_hashes(h, rh):
    """(zh, zl) per column: XOR of its rows' hashes (cdist/search.py:88-105)."""
    zh = np.zeros(len(h), dtype=np.uint64)
    zl = np.zeros(len(h), dtype=np.uint64)
    cache = {}
    for j, a in enumerate(h):
        v = cache.get(a)
        if v is None:
            x = y = 0
            m = a
            while m:
                low = m & -m
                r = low.bit_length() - 1
                x ^= rh[0][r]
                y ^= rh[1][r]
                m ^= low
            v = cache[a] = (x, y)
        zh[j], zl[j] = v
    return zh, zl


def table_log2slots(ncols, k, load=0.5):
    """log2 of the slot count for all subsets of size <= k at the given load (cdist/search.py:122-125)."""
    total = sum(math.comb(ncols, i) for i in range(k + 1))
    return max(10, math.ceil(math.log2(total / load + 1)))


class Table:
    """All column subsets of size <= k of a problem in a device hash table (cdist/search.py:108-142)."""

    def __init__(self, p, k, seed=12345, load=0.5, labelfree=False):
        assert labelfree or p.nobs <= 64, "more than 64 observables need a label-free table"
        self.problem, self.k, self.labelfree = p, k, labelfree
        self.wide = p.nobs > 16 and not labelfree
        self.rh = row_hashes(p.nrows, seed)
        self.zh, self.zl = column_hashes(p.h, self.rh)
        self.lg = np.zeros(p.ncols, dtype=np.uint64) if labelfree else np.array(p.l, dtype=np.uint64)
        self.log2slots = table_log2slots(p.ncols, k, load)
        L = _lib.lib()
        self.L = L
        def create():
            h = L.tdp_tab_create(self.log2slots, 1 if self.wide else 0)
            if not h:
                raise MemoryError("table with 2^%d slots: %s" % (self.log2slots, L.tdp_last_error().decode()))
            return h

        self.handle = _lib.retry_after_release(create)
        t = time.time()
        _lib.check(L.tdp_tab_build(self.handle, p.ncols, self.zh.ctypes.data, self.zl.ctypes.data,
                                   self.lg.ctypes.data, k), "tdp_tab_build")
        self.build_seconds = time.time() - t
        self.entries = L.tdp_tab_count(self.handle)

    def hashes_for(self, h):
        """Hashes of arbitrary syndromes of this table's problem."""
        return column_hashes(h, self.rh)

    def lookup(self, zh, zl, lg, R):
        """Batched lookups: arrays (or scalars) of hashes, labels and budgets -> int array of costs (-1 = none)."""
        t = _lib.torch()
        zh = np.atleast_1d(np.asarray(zh, dtype=np.uint64))
        m = zh.size
        zl = np.broadcast_to(np.asarray(zl, dtype=np.uint64), (m,))
        lg = np.broadcast_to(np.asarray(lg, dtype=np.uint64), (m,))
        R = np.broadcast_to(np.asarray(R, dtype=np.int32), (m,))
        dev = [t.from_numpy(np.ascontiguousarray(a).view(np.int64 if a.dtype == np.uint64 else np.int32)).cuda()
               for a in (zh, zl, lg, R)]
        out = t.empty(m, dtype=t.int32, device="cuda")
        _lib.check(self.L.tdp_tab_lookup_batch(self.handle, dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(),
                                               dev[3].data_ptr(), m, out.data_ptr(), _lib.stream_handle()),
                   "tdp_tab_lookup_batch")
        return out.cpu().numpy()

    def free(self):
        if self.handle:
            self.L.tdp_tab_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
