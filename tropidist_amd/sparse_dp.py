"""Weight-truncated sparse Rank DP with meet in the middle, tables resident on the GPU.

Call surface of the reference's cdist/sparse_dp.py: Trellis(h, lrow, order, nrows) / Trellis.for_query, BACKENDS,
Sweep, query(tr, K, mode, split, ...), distance(p, ...), choose_split / choose_schedule, BudgetExceeded, summary.

The semantics are the reference's (module docstring of cdist/sparse_dp.py:1-30):
  truncation  a state of weight > K is dropped (its value only grows along the sweep);
  lookahead   a state with weight + h > K is dropped, h an admissible covering bound (use_h);
  mitm        a backward sweep keeps weights <= K - a(c), a forward sweep keeps weights <= a(c) - 1 and joins its
              new or improved states (weights <= a(c)) with the backward table at every cut;
  witness     a traceback through tables recomputed from checkpoints every isqrt(N) cuts.

Backend "hip" (HipTables): a table is a device array of 16-byte states (keys of up to 120 row slots, the value in
the top byte) with a device hash index; one column step is one kernel pass over the old states (tdp_sp_step), the
join is a batched probe with a deterministic minimum (tdp_sp_join).  Tables never leave the device except for
statistics and single-key lookups.  The dense fallback of the reference (Sweep(dense=...)) is not provided by this
backend.
"""
from __future__ import annotations

import ctypes
import heapq
import math
import time

import numpy as np

from . import _lib, trellis

INF = 255
MAX_WEIGHT = INF - 1
Q = 720720
KEYMASK = (1 << 56) - 1
MAX_SLOTS = 120                 # 64 + 56 key bits of a device state
U64 = np.uint64
M64 = (1 << 64) - 1


def _bits(x):
    out = []
    while x:
        low = x & -x
        out.append(low.bit_length() - 1)
        x ^= low
    return out


class Trellis:
    """Minimal-span trellis of one query min{|e| : H e = 0, (L e)_obs = 1} in a column order, with rows on fixed
    bit positions ("slots") so that a state means the same in both sweep directions (cdist/sparse_dp.py:63-145).
    Cut c separates positions [0, c) from [c, N)."""

    def __init__(self, h, lrow, order, nrows, name=""):
        self.name = name
        self.order = list(order)
        self.N = N = len(self.order)
        hrows = trellis.minimal_span_form(trellis.rows_from_columns(h, self.order, nrows))
        lr = trellis.reduce_target_row(lrow, hrows)
        self.feasible = lr != 0
        if not self.feasible:
            return
        rows = list(hrows) + [lr]
        self.rows = rows
        self.target = [0] * len(hrows) + [1]
        self.L = len(rows) - 1
        self.start = [trellis.lowbit(r) for r in rows]
        self.end = [r.bit_length() - 1 for r in rows]
        # interval colouring of the closed spans [start, end]: lowest free slot first
        slot = [0] * len(rows)
        free, busy, nslots = [], [], 0
        for r in sorted(range(len(rows)), key=lambda r: self.start[r]):
            while busy and busy[0][0] < self.start[r]:
                heapq.heappush(free, heapq.heappop(busy)[1])
            if free:
                s = heapq.heappop(free)
            else:
                s, nslots = nslots, nslots + 1
            slot[r] = s
            heapq.heappush(busy, (self.end[r], s))
        if nslots > MAX_SLOTS:
            raise ValueError("trellis needs %d bits per state (at most %d supported)" % (nslots, MAX_SLOTS))
        self.slot, self.nslots = slot, nslots
        col = [0] * N
        for r, row in enumerate(rows):
            for pos in _bits(row):
                col[pos] |= 1 << slot[r]
        self.col = col
        fsm, fsb, bsm, bsb = [0] * N, [0] * N, [0] * N, [0] * N
        starts_at, ends_at = [[] for _ in range(N)], [[] for _ in range(N)]
        for r in range(len(rows)):
            b = 1 << slot[r]
            fsm[self.end[r]] |= b
            bsm[self.start[r]] |= b
            if self.target[r]:
                fsb[self.end[r]] |= b
                bsb[self.start[r]] |= b
            starts_at[self.start[r]].append((slot[r], self.target[r]))
            ends_at[self.end[r]].append((slot[r], self.target[r]))
        self.fsel_mask, self.fsel_bits, self.bsel_mask, self.bsel_bits = fsm, fsb, bsm, bsb
        self.starts_at, self.ends_at = starts_at, ends_at
        delta = [[] for _ in range(N + 1)]
        for r in range(len(rows)):
            delta[self.start[r] + 1].append(r)
            delta[self.end[r] + 1].append(~r)
        act, cur = [], set()
        for c in range(N + 1):
            for x in delta[c]:
                if x >= 0:
                    cur.add(x)
                else:
                    cur.discard(~x)
            act.append(sorted(cur))
        self.active_rows = act
        self.active = [sum(1 << slot[r] for r in a) for a in act]
        self.width = [len(a) for a in act]
        Ls = 1 << slot[self.L]
        self.tmask = [Ls if self.L in set(a) else 0 for a in act]
        self._bd = {}

    @classmethod
    def for_query(cls, p, obs, order):
        lrow = 0
        for pos, j in enumerate(order):
            if (p.l[j] >> obs) & 1:
                lrow |= 1 << pos
        return cls(p.h, lrow, order, p.nrows, name="%s obs %d" % (getattr(p, "name", ""), obs))

    def _bound_data(self, forward):
        """Per cut: (masks, weights, extra) of h(s) = ceil((extra + sum_k w_k |(s ^ tmask) & mask_k|) / Q)
        (cdist/sparse_dp.py:148-192): row r gets the weight 1 / max over remaining columns j containing r of
        |supp(j) /\\ R_c|, R_c the active rows plus the target row while it lies wholly ahead, rounded down to 1/Q."""
        N, rows = self.N, self.rows
        S = np.zeros((len(rows), N), dtype=np.int32)
        for r, row in enumerate(rows):
            S[r, _bits(row)] = 1
        out = []
        for c in range(N + 1):
            rel = list(self.active_rows[c])
            lone = False
            if forward and self.start[self.L] >= c:
                rel.append(self.L)
                lone = True
            if not forward and self.end[self.L] < c:
                rel.append(self.L)
                lone = True
            if not rel:
                out.append(((), (), 0))
                continue
            lo, hi = (c, max(self.end[r] for r in rel) + 1) if forward else (min(self.start[r] for r in rel), c)
            sub = S[rel, lo:hi]
            cmax = (sub * sub.sum(axis=0)).max(axis=1)
            assert (cmax >= 1).all()
            w = Q // cmax
            masks, extra = {}, 0
            for r, wr in zip(rel, w.tolist()):
                if lone and r == self.L:
                    extra = wr
                    continue
                masks[wr] = masks.get(wr, 0) | (1 << self.slot[r])
            ks = sorted(masks)
            out.append((tuple(masks[k] for k in ks), tuple(int(k) for k in ks), int(extra)))
        return out

    def bound_data(self, forward):
        if forward not in self._bd:
            self._bd[forward] = self._bound_data(forward)
        return self._bd[forward]


# ---------------------------------------------------------------------------------------------- device tables
def _split(x):
    return (ctypes.c_uint64 * 2)(x & M64, x >> 64)


class DevTable:
    """n states on the device (16 bytes each) and, once built, their hash index."""
    __slots__ = ("buf", "n", "slots", "capacity")

    def __init__(self, buf, n):
        self.buf, self.n, self.slots, self.capacity = buf, n, None, 0

    def ptr(self):
        return self.buf.data_ptr()


class HipTables:
    """Sparse tables on the GPU (the backend interface of cdist/sparse_dp.py:300-445 / 506-605)."""
    name = "hip"

    def __init__(self, tr, stream=None):
        if tr.nslots > MAX_SLOTS:
            raise ValueError("the device tables hold at most %d row slots (this trellis needs %d)"
                             % (MAX_SLOTS, tr.nslots))
        self.tr = tr
        self.L = _lib.lib()
        self.stream = stream
        self.counters = _lib.device_empty(64)
        self._host_counters = np.zeros(8, dtype=np.uint64)
        self._bounds = {}
        self._sh = _lib.stream_handle(stream)       # the query's stream, fixed for the table set's life

    def sh(self):
        return self._sh

    def _alloc(self, nstates):
        return _lib.device_empty(16 * max(1, nstates))

    def _sync_counters(self, k):
        _lib.check(self.L.tdp_read_back(self._host_counters.ctypes.data_as(ctypes.c_void_p), self.counters.data_ptr(),
                                        8 * k, self._sh), "tdp_read_back")
        return self._host_counters[:k].tolist()

    def index(self, tab):
        if tab.slots is None:
            cap = 64
            while cap < 2 * tab.n:
                cap <<= 1
            tab.slots = _lib.device_empty(4 * cap)
            tab.capacity = cap
            _lib.check(self.L.tdp_sp_index(tab.ptr(), tab.n, tab.slots.data_ptr(), cap, self.sh()), "tdp_sp_index")
        return tab

    def start(self):
        return self.from_arrays([0], np.zeros(1, dtype=np.uint8))

    @staticmethod
    def size(tab):
        return tab.n

    @staticmethod
    def nbytes(tab):
        return 16 * tab.n

    def arrays(self, tab):
        """(keys, vals) of a table: keys as Python ints (object array) when any exceeds 64 bits, else uint64."""
        raw = tab.buf[:16 * tab.n].cpu().numpy().view(np.uint64).reshape(-1, 2) if tab.n else np.zeros((0, 2), U64)
        vals = (raw[:, 1] >> U64(56)).astype(np.uint8)
        hi = raw[:, 1] & U64(KEYMASK)
        if tab.n and hi.any():
            keys = np.array([int(a) | (int(b) << 64) for a, b in zip(raw[:, 0], hi)], dtype=object)
        else:
            keys = raw[:, 0].copy()
        return keys, vals

    def from_arrays(self, keys, vals):
        n = len(vals)
        raw = np.zeros((max(1, n), 2), dtype=U64)
        for i, (k, v) in enumerate(zip(keys, vals)):
            k = int(k)
            raw[i, 0] = k & M64
            raw[i, 1] = (k >> 64) | (int(v) << 56)
        buf = self._alloc(n)
        t = _lib.torch()
        buf.copy_(t.from_numpy(raw.view(np.uint8).reshape(-1)[:16 * max(1, n)]))
        return DevTable(buf, n)

    def lookup_many(self, tab, keys):
        """Values of the given states (INF where absent)."""
        m = len(keys)
        if not m or not tab.n:
            return np.full(m, INF, dtype=np.uint8)
        self.index(tab)
        q = np.zeros((m, 2), dtype=U64)
        for i, k in enumerate(keys):
            k = int(k)
            q[i] = (k & M64, k >> 64)
        t = _lib.torch()
        dq = t.from_numpy(q.view(np.uint8).reshape(-1)).cuda()
        dv = _lib.device_empty(m)
        _lib.check(self.L.tdp_sp_lookup(tab.ptr(), tab.slots.data_ptr(), tab.capacity, dq.data_ptr(), m,
                                        dv.data_ptr(), self.sh()), "tdp_sp_lookup")
        return dv[:m].cpu().numpy()

    def lookup(self, tab, s):
        return int(self.lookup_many(tab, [s])[0])

    def bound(self, forward):
        """Per-cut lookahead data on the device: (mask tensor, weight tensor, ncls, extra)."""
        if forward not in self._bounds:
            t = _lib.torch()
            per = []
            for masks, weights, extra in self.tr.bound_data(forward):
                m = np.zeros((max(1, len(masks)), 2), dtype=U64)
                for i, x in enumerate(masks):
                    m[i] = (x & M64, x >> 64)
                w = np.zeros(max(1, len(weights)), dtype=np.int64)
                w[:len(weights)] = weights
                per.append((t.from_numpy(m.reshape(-1)).cuda(), t.from_numpy(w).cuda(), len(masks), extra))
            self._bounds[forward] = per
        return self._bounds[forward]

    def step(self, tab, upd, sel_mask, sel_bits, cap, K, tmask, hd, keep_below=None, want_changed=False):
        """One column: T'[s] = min(T[s], T[s ^ upd] + 1), the closing-row filter, the cap and the lookahead
        (cdist/sparse_dp.py:338-403); returns (table, changed)."""
        self.index(tab)
        n = tab.n
        out = self._alloc(2 * n)
        ch = self._alloc(2 * n) if want_changed else None
        if hd is None:
            ncls, mp, wp, extra, use_h = 0, None, None, 0, 0
        else:
            dm, dw, ncls, extra = hd
            mp, wp, use_h = dm.data_ptr(), dw.data_ptr(), 1
        rc = self.L.tdp_sp_step(tab.ptr(), n, tab.slots.data_ptr(), tab.capacity, _split(upd), _split(sel_mask),
                                _split(sel_bits), cap, keep_below or 0, ncls, mp, wp, extra, _split(tmask), K, use_h,
                                out.data_ptr(), 2 * n, ch.data_ptr() if ch is not None else None, 2 * n,
                                self.counters.data_ptr(), self.sh())
        _lib.check(rc, "tdp_sp_step")
        no, nc, overflow = self._sync_counters(3)
        if overflow:
            raise RuntimeError("sparse step output overflow")
        return DevTable(out, no), (DevTable(ch, nc) if want_changed else None)

    @staticmethod
    def changed_size(ch):
        return ch.n

    def persistent(self, tab):
        """A table kept across steps: its states in a buffer of its own size."""
        if tab.buf.numel() > 16 * max(1, tab.n) * 5 // 4 + 16:
            buf = self._alloc(tab.n)
            buf[:16 * tab.n].copy_(tab.buf[:16 * tab.n])
            return DevTable(buf, tab.n)
        return tab

    def truncate(self, tab, below):
        out = self._alloc(tab.n)
        _lib.check(self.L.tdp_sp_truncate(tab.ptr(), tab.n, below, out.data_ptr(), self.counters.data_ptr(),
                                          self.sh()), "tdp_sp_truncate")
        return DevTable(out, self._sync_counters(1)[0])

    def join(self, ch, btab, tmask):
        """Least value(f) + value(b) over changed forward states f and backward states b = f ^ tmask:
        (total, forward key, forward value, backward key) or None (cdist/sparse_dp.c:204-236); ties go to the
        least forward key."""
        if not ch.n or not btab.n:
            return None
        self.index(btab)
        _lib.check(self.L.tdp_sp_join(ch.ptr(), ch.n, btab.ptr(), btab.slots.data_ptr(), btab.capacity,
                                      _split(tmask), self.counters.data_ptr(), self.sh()), "tdp_sp_join")
        tot, hi, lo, fval = self._sync_counters(4)
        if tot == M64:
            return None
        fkey = lo | (hi << 64)
        return int(tot), fkey, int(fval), fkey ^ tmask

    def hist(self, tab):
        h = _lib.device_empty(8 * 256)
        _lib.check(self.L.tdp_sp_hist(tab.ptr(), tab.n, h.data_ptr(), self.sh()), "tdp_sp_hist")
        return h.cpu().numpy().view(np.uint64).astype(np.int64)


BACKENDS = {"hip": HipTables, "c": HipTables}


class BudgetExceeded(RuntimeError):
    """A table outgrew max_states (kind "states") or a sweep ran past its deadline (kind "deadline")."""

    def __init__(self, message, kind):
        super().__init__(message)
        self.kind = kind


def _make_backend(backend, tr):
    if isinstance(backend, str):
        return BACKENDS[backend](tr)
    return backend


# ---------------------------------------------------------------------------------------------- sweeps
class Sweep:
    """One direction of the truncated sweep, with tables kept at checkpoints (cdist/sparse_dp.py:621-820).

    direction +1: cuts 0 -> N from {0: 0}; -1: cuts N -> 0.  caps: one per cut, non-increasing in the sweep
    direction.  frontier (the forward half of the meet in the middle): the step out of cut p keeps weights
    <= caps[p] - 1 and returns the new or improved states of weight <= caps[p] as `changed`; where caps drops,
    the table before the cut is returned as `extra` to be joined in full."""

    def __init__(self, tr, direction, K, cap=None, use_h=True, every=None, frontier=False, backend="hip",
                 dense=None, dense_max=1 << 30, max_states=None, deadline=None, caps=None):
        self.tr, self.direction, self.K = tr, direction, K
        self.caps = [min(int(x), K) for x in caps] if caps is not None else [min(cap, K)] * (tr.N + 1)
        if max(self.caps) > MAX_WEIGHT:
            raise ValueError("a sweep keeping weights up to %d: the tables hold at most %d (8-bit values)"
                             % (max(self.caps), MAX_WEIGHT))
        if direction > 0:
            assert all(x >= y for x, y in zip(self.caps, self.caps[1:])), "forward caps must not increase"
        else:
            assert all(x <= y for x, y in zip(self.caps, self.caps[1:])), "backward caps must not decrease"
        if dense is not None and not frontier:
            raise ValueError("the GPU tables have no dense fallback (dense=None)")
        self.use_h, self.frontier = use_h, frontier
        self.be = _make_backend(backend, tr)
        self.hd = self.be.bound(direction > 0) if use_h else None
        self.every = every or max(1, int(math.isqrt(tr.N)))
        self.max_states, self.deadline = max_states, deadline
        self.ckpt = {}
        self.counts = [0] * (tr.N + 1)
        self.changed_counts = [0] * (tr.N + 1)
        self.dense_cuts = 0
        self.peak = 0
        self.peak_bytes = 0
        self.steps_done = 0

    def first(self):
        return (0 if self.direction > 0 else self.tr.N), self.be.start()

    def _step(self, tab, p, want_changed):
        tr, fwd = self.tr, self.direction > 0
        cut = p + 1 if fwd else p
        upd = tr.col[p]
        sm, sb = (tr.fsel_mask[p], tr.fsel_bits[p]) if fwd else (tr.bsel_mask[p], tr.bsel_bits[p])
        self.steps_done += 1
        cap, kb = self.caps[cut], None
        if self.frontier:
            cap = kb = self.caps[p]
        tab, ch = self.be.step(tab, upd, sm, sb, cap, self.K, tr.tmask[cut], self.hd[cut] if self.hd else None,
                               kb, want_changed)
        if self.frontier and self.caps[cut] < cap:
            return cut, self.be.truncate(tab, self.caps[cut]), ch, tab
        return cut, tab, ch, None

    def _iter(self, cut, tab, stop, want_changed=False):
        rng = range(cut, stop) if self.direction > 0 else range(cut - 1, stop - 1, -1)
        for p in rng:
            c, tab, ch, extra = self._step(tab, p, want_changed)
            yield c, tab, ch, extra

    def is_ckpt(self, c):
        return c % self.every == 0 or c == 0 or c == self.tr.N

    def run(self, on_cut=None):
        """Full sweep keeping checkpoints and per-cut counts; on_cut(cut, table, changed, extra) may return True to
        stop.  Returns the last cut reached."""
        c0, tab = self.first()
        self.ckpt[c0] = tab
        self.counts[c0] = self.changed_counts[c0] = 1
        self.peak = 1
        if on_cut and on_cut(c0, tab, tab, None):
            return c0
        stop = self.tr.N if self.direction > 0 else 0
        ck_bytes = self.be.nbytes(tab)
        for c, tab, ch, extra in self._iter(c0, tab, stop, want_changed=on_cut is not None):
            n = self.be.size(tab)
            nc = self.be.changed_size(ch) if ch is not None else 0
            self.counts[c], self.changed_counts[c] = n, nc
            self.peak = max(self.peak, n, nc)
            if self.is_ckpt(c):
                self.ckpt[c] = self.be.persistent(tab)
                ck_bytes += self.be.nbytes(tab)
            self.peak_bytes = max(self.peak_bytes, ck_bytes + self.be.nbytes(tab))
            if self.max_states is not None and max(n, nc) > self.max_states:
                raise BudgetExceeded("%d states at cut %d" % (max(n, nc), c), "states")
            if self.deadline is not None and time.monotonic() > self.deadline:
                raise BudgetExceeded("deadline at cut %d" % c, "deadline")
            if on_cut and on_cut(c, tab, ch, extra):
                return c
        return stop

    def segment(self, lo, hi):
        """Tables at the cuts lo..hi, recomputed from the nearest checkpoint: dict cut -> table."""
        out = {}
        if self.direction > 0:
            c0 = max(c for c in self.ckpt if c <= lo)
            tab = self.ckpt[c0]
            if c0 >= lo:
                out[c0] = tab
            for c, tab, _, _ in self._iter(c0, tab, hi):
                if c >= lo:
                    out[c] = self.be.persistent(tab)
        else:
            c0 = min(c for c in self.ckpt if c >= hi)
            tab = self.ckpt[c0]
            if c0 <= hi:
                out[c0] = tab
            for c, tab, _, _ in self._iter(c0, tab, lo):
                if c <= hi:
                    out[c] = self.be.persistent(tab)
        return out

    def table(self, c, cache):
        """Table at cut c for cuts requested against the sweep's direction (cache reused between calls)."""
        if c not in cache:
            cache.clear()
            if self.direction < 0:
                cache.update(self.segment(c, min(x for x in self.ckpt if x >= c)))
            else:
                cache.update(self.segment(max(x for x in self.ckpt if x <= c), c))
        return cache[c]

    def traceback(self, cut, state, value):
        """Columns (problem indices) of a partial solution reaching `state` with weight `value` at `cut`
        (cdist/sparse_dp.py:788-819)."""
        tr, used = self.tr, []
        s, v = int(state), int(value)
        if self.direction > 0:
            c = cut
            while c > 0:
                lo = max(0, (c - 1) // self.every * self.every)
                tabs = self.segment(lo, c - 1)
                for p in range(c - 1, lo - 1, -1):
                    s, v = self._undo(tabs[p], s, v, s | tr.fsel_bits[p], tr.col[p], tr.order[p], used)
                c = lo
        else:
            c = cut
            while c < tr.N:
                hi = min(tr.N, (c // self.every + 1) * self.every)
                tabs = self.segment(c + 1, hi)
                for p in range(c, hi):
                    s, v = self._undo(tabs[p + 1], s, v, s | tr.bsel_bits[p], tr.col[p], tr.order[p], used)
                c = hi
        assert s == 0 and v == 0, (s, v)
        return used

    def _undo(self, tab, s, v, s_mid, upd, column, used):
        stay, move = self.be.lookup_many(tab, [s_mid, s_mid ^ upd]) if upd else (self.be.lookup(tab, s_mid), INF)
        if int(stay) == v:
            return s_mid, v
        assert upd and int(move) == v - 1, "traceback lost the path"
        used.append(column)
        return s_mid ^ upd, v - 1


# ---------------------------------------------------------------------------------------------- the split
def _values(be, tab):
    return be.arrays(tab)[1]


def weight_profile(tr, direction, cap, backend="hip"):
    """States of each weight <= cap summed over all cuts, for a sweep keeping weights <= cap (no lookahead)."""
    sw = Sweep(tr, direction, cap, cap, use_h=False, backend=backend)
    c0, tab = sw.first()
    tot = np.zeros(cap + 1)
    tot[0] = 1
    for c, tab, _, _ in sw._iter(c0, tab, tr.N if direction > 0 else 0):
        tot += sw.be.hist(tab)[:cap + 1]
    return tot


def choose_split(tr, K, backend="hip", probe=3, budget=2e9):
    """Split a minimising sum_{w<=a-1} F(w) + F(a-1) + 2 sum_{w<=K-a} B(w), F and B measured to weight `probe` and
    extrapolated with the growth of their last two weights (cdist/sparse_dp.py:839-867).  Returns (a, info)."""
    t0 = time.time()
    prof = {}
    for direction in (+1, -1):
        p = weight_profile(tr, direction, min(2, probe), backend)
        if probe > 2 and len(p) > 2 and p[1] > 0 and p[2] * p[2] / p[1] < budget:
            p = weight_profile(tr, direction, probe, backend)
        prof[direction] = p

    def extend(p, upto):
        p = list(p)
        r = p[-1] / p[-2] if len(p) > 1 and p[-2] > 0 else 1.0
        while len(p) <= upto:
            p.append(p[-1] * max(r, 1.0))
        return p

    F, B = extend(prof[+1], K), extend(prof[-1], K)
    costs = {a: sum(F[:a]) + F[a - 1] + 2 * sum(B[:K - a + 1]) for a in range(1, K + 1)}
    a = min(costs, key=costs.get)
    return a, {"forward_profile": [float(x) for x in prof[+1]], "backward_profile": [float(x) for x in prof[-1]],
               "predicted_work": {int(k): float(v) for k, v in costs.items()}, "probe_seconds": time.time() - t0}


def cut_profiles(tr, direction, cap, backend="hip"):
    """States of each weight <= cap at every cut (array (N + 1, cap + 1))."""
    sw = Sweep(tr, direction, cap, cap, use_h=False, backend=backend)
    c0, tab = sw.first()
    out = np.zeros((tr.N + 1, cap + 1))
    out[c0, 0] = 1
    for c, tab, _, _ in sw._iter(c0, tab, tr.N if direction > 0 else 0):
        out[c] = sw.be.hist(tab)[:cap + 1]
    return out


def probe_profiles(tr, backend="hip", probe=3, budget=2e9):
    """Per-cut state counts per weight in both directions up to weight `probe` (or 2 where weight 3 would exceed
    the budget), kept on the trellis; returns (profiles, seconds spent now) (cdist/sparse_dp.py:882-903)."""
    t0 = time.time()

    def measured(direction, cap):
        key = ("profile", direction, cap)
        if key not in tr._bd:
            tr._bd[key] = cut_profiles(tr, direction, cap, backend)
        return tr._bd[key]

    prof = {}
    for direction in (+1, -1):
        b = budget[direction] if isinstance(budget, dict) else budget
        p = measured(direction, min(2, probe))
        tot = p.sum(axis=0)
        if probe > 2 and tot[1] > 0 and tot[2] * tot[2] / tot[1] < b:
            p = measured(direction, probe)
        prof[direction] = p
    return prof, time.time() - t0


def choose_schedule(tr, K, backend="hip", probe=3, budget=2e9, window=100):
    """A non-increasing split a(c) per cut minimising sum_c F_c(<= a-1) + F_c(a-1) + 2 B_c(<= K-a) by a dynamic
    programme over cuts and values of a (cdist/sparse_dp.py:906-981).  Returns (schedule, info)."""
    prof, probe_seconds = probe_profiles(tr, backend, probe, budget)

    def wsum(x):
        cs = np.concatenate([[0.0], np.cumsum(x)])
        idx = np.arange(x.size)
        return cs[np.minimum(x.size, idx + window + 1)] - cs[np.maximum(0, idx - window)]

    def extend(p):
        top = p.shape[1] - 1
        if top >= K:
            return p[:, :K + 1]
        hi = wsum(p[:, top])
        lo = wsum(p[:, top - 1]) if top >= 1 else hi
        r = np.maximum(np.where(lo > 0, hi / np.maximum(lo, 1e-300), 1.0), 1.0)
        cols = [p[:, w] for w in range(top + 1)]
        while len(cols) <= K:
            cols.append(cols[-1] * r)
        return np.stack(cols, axis=1)

    F, B = extend(prof[+1]), extend(prof[-1])
    cF, cB = np.cumsum(F, axis=1), np.cumsum(B, axis=1)
    n = tr.N + 1
    cost = np.zeros((n, K + 1))
    for a in range(1, K + 1):
        cost[:, a] = cF[:, a - 1] + F[:, a - 1] + 2 * cB[:, K - a]
    cost[:, 0] = np.inf
    best = np.empty((n, K + 1))
    arg = np.zeros((n, K + 1), dtype=np.int64)
    best[0] = cost[0]
    for c in range(1, n):
        prev = best[c - 1]
        suf = np.minimum.accumulate(prev[::-1])[::-1]
        sarg = np.empty(K + 1, dtype=np.int64)
        m, ai = np.inf, K
        for a in range(K, -1, -1):
            if prev[a] < m:
                m, ai = prev[a], a
            sarg[a] = ai
        best[c] = cost[c] + suf
        arg[c] = sarg
    a = int(np.argmin(best[n - 1]))
    sched = [0] * n
    for c in range(n - 1, -1, -1):
        sched[c] = a
        if c:
            a = int(arg[c, a])
    idx = np.arange(n)
    av = np.array(sched)
    fwd = cF[idx, av - 1] + F[idx, av - 1]
    bwd = cB[idx, K - av]
    every = max(1, int(math.isqrt(tr.N)))
    ck = (idx % every == 0) | (idx == 0) | (idx == n - 1)
    seg = np.add.reduceat(bwd, np.arange(0, n, every)).max()
    states = bwd[ck].sum() + seg + cF[idx, av - 1][ck].sum() + 4 * max(fwd.max(), bwd.max())
    uniform = {a: float(cost[:, a].sum()) for a in range(1, K + 1)}
    return sched, {"predicted_work": float(best[n - 1].min()), "uniform_predicted_work": uniform,
                   "predicted_peak_states": float(max(fwd.max(), bwd.max())),
                   "predicted_mean_states": float((fwd.sum() + bwd.sum()) / (2 * n)),
                   "predicted_bytes": float(16 * states),
                   "schedule": _runs(sched), "forward_profile": prof[+1].sum(axis=0).tolist(),
                   "backward_profile": prof[-1].sum(axis=0).tolist(), "probe_seconds": probe_seconds}


# ---------------------------------------------------------------------------------------------- one query
def query(tr, K, mode="mitm", split=None, use_h=True, witness=True, lower=0, every=None, backend="hip",
          dense=None, dense_max=1 << 30, log=None, max_states=None, deadline=None):
    """Least weight of a solution of the trellis' query if it is <= K (else None), with per-cut statistics and a
    witness (cdist/sparse_dp.py:985-1095).  mode "forward": one sweep keeping weights <= K.  mode "mitm": split a(c)
    per cut, non-increasing (an int, "auto", "schedule" or a list of N + 1 values); the backward sweep keeps
    weights <= K - a(c), the forward frontier weights <= a(c) - 1 and its new or improved states are joined with the
    backward table at every cut.  See the reference's docstring for why every solution of weight <= K is seen."""
    t0 = time.time()
    bname = backend if isinstance(backend, str) else backend.name
    rec = {"K": K, "mode": mode, "use_h": bool(use_h), "value": None, "witness": None, "width": max(tr.width),
           "backend": bname}

    def new_be():
        return BACKENDS[backend](tr) if isinstance(backend, str) else backend.__class__(tr)

    be = new_be()
    if mode == "forward":
        fw = Sweep(tr, +1, K, K, use_h, every, backend=be, dense=dense, dense_max=dense_max, max_states=max_states,
                   deadline=deadline)
        fw.run()
        tab = fw.ckpt[tr.N]
        rec.update(forward_counts=fw.counts, peak_states=fw.peak, peak_table_bytes=fw.peak_bytes,
                   sweep_seconds=time.time() - t0, steps=fw.steps_done, dense_cuts=fw.dense_cuts)
        v = be.lookup(tab, 0)
        if v != INF:
            rec["value"] = v
            if witness:
                t1 = time.time()
                rec["witness"] = sorted(fw.traceback(tr.N, 0, v))
                rec["traceback_seconds"] = time.time() - t1
        rec["seconds"] = time.time() - t0
        return rec
    predicted = None
    if split == "auto":
        split, rec["split_choice"] = choose_split(tr, K, new_be())
        predicted = rec["split_choice"]["predicted_work"][split]
    elif split == "schedule":
        split, rec["split_choice"] = choose_schedule(tr, K, new_be())
        predicted = rec["split_choice"]["predicted_work"]
    if split is None:
        split = K // 2 + 1
    sched = [max(1, min(int(x), K)) for x in split] if isinstance(split, (list, tuple)) else \
        [max(1, min(int(split), K))] * (tr.N + 1)
    assert len(sched) == tr.N + 1 and all(x >= y for x, y in zip(sched, sched[1:])), "split must not increase"
    bw = Sweep(tr, -1, K, use_h=use_h, every=every, backend=be, dense=dense, dense_max=dense_max,
               max_states=max_states, deadline=deadline, caps=[K - x for x in sched])
    fw = Sweep(tr, +1, K, use_h=use_h, every=bw.every, frontier=True, backend=new_be(), max_states=max_states,
               deadline=deadline, caps=sched)
    if log:
        log("K=%d split a=%s: backward sweep (weights <= %s)%s" % (
            K, _runs(sched), _runs([K - x for x in sched]), "" if predicted is None else
            ", predicted work %.3g" % predicted))
    bw.run()
    t1 = time.time()
    if log:
        log("backward sweep done: %.1f s, peak %d states, %.3g states in all" % (t1 - t0, bw.peak, sum(bw.counts)))
    best, cache, joins = [None], {}, [0]

    def join(c, tab, ch, extra):
        for part in (ch, extra):
            if part is None or not be.changed_size(part):
                continue
            joins[0] += be.changed_size(part)
            r = be.join(part, bw.table(c, cache), tr.tmask[c])
            if r is not None and (best[0] is None or r[0] < best[0][0]):
                best[0] = (r[0], c) + tuple(r[1:])
        return best[0] is not None and best[0][0] <= max(lower, 0)

    fw.run(on_cut=join)
    cache.clear()
    rec.update(forward_counts=fw.counts, backward_counts=bw.counts, changed_counts=fw.changed_counts,
               peak_states=max(fw.peak, bw.peak), peak_table_bytes=fw.peak_bytes + bw.peak_bytes,
               backward_seconds=t1 - t0, forward_seconds=time.time() - t1,
               split=sched[0] if len(set(sched)) == 1 else _runs(sched), split_values=sched, joined=joins[0],
               steps=fw.steps_done + bw.steps_done, dense_cuts=bw.dense_cuts)
    if best[0] is not None and best[0][0] <= K:
        value, c, sl, fl, sr = best[0]
        rec["value"], rec["join_cut"] = value, c
        if witness:
            t2 = time.time()
            rec["witness"] = sorted(fw.traceback(c, sl, fl) + bw.traceback(c, sr, value - fl))
            rec["traceback_seconds"] = time.time() - t2
    rec["seconds"] = time.time() - t0
    return rec


def _runs(x):
    """A per-cut list as 'value' or 'value@cut,...' where it changes."""
    out = [(v, c) for c, v in enumerate(x) if c == 0 or v != x[c - 1]]
    return str(out[0][0]) if len(out) == 1 else ",".join("%d@%d" % vc for vc in out)


# ---------------------------------------------------------------------------------------------- distance
def default_order(p, seeds=2):
    from . import auto
    return auto.best_order(p)[0]


def distance(p, order=None, start=1, kmax=64, mode="mitm", use_h=True, witness=True, log=None, every=None,
             backend="hip", dense=None, split=None):
    """Exact distance by iterative deepening of K from `start`, a lower bound (cdist/sparse_dp.py:1110-1151)."""
    t0 = time.time()
    order = default_order(p) if order is None else order
    trs = [Trellis.for_query(p, i, order) for i in range(p.nobs)]
    out = {"distance": None, "witness": None, "runs": [], "lower_bound": start,
           "order_cut_rank": max((max(t.width) for t in trs if t.feasible), default=0)}
    for K in range(start, kmax + 1):
        best = None
        for i, tr in enumerate(trs):
            if not tr.feasible:
                continue
            r = query(tr, K, mode=mode, use_h=use_h, witness=witness, lower=K, every=every, backend=backend,
                      dense=dense, split=split(K) if callable(split) else split)
            r["obs"] = i
            out["runs"].append(summary(r))
            if log:
                log("  K=%d obs %d: value %s peak %d states %.2fs" % (K, i, r["value"], r["peak_states"],
                                                                       r["seconds"]))
            if r["value"] is not None and (best is None or r["value"] < best["value"]):
                best = r
                if r["value"] <= K:
                    break
        if best is not None:
            out["distance"] = best["value"]
            if witness:
                cols = best["witness"]
                s = lg = 0
                for j in cols:
                    s ^= p.h[j]
                    lg ^= p.l[j]
                assert s == 0 and lg != 0 and len(cols) == best["value"] == len(set(cols)), "witness check failed"
                out["witness"] = cols
            break
        out["lower_bound"] = K + 1
    out["seconds"] = time.time() - t0
    return out


def summary(r):
    """A query record without the per-cut arrays (their totals instead)."""
    out = {k: v for k, v in r.items() if not k.endswith("_counts") and k not in ("witness", "split_values")}
    for k in ("forward_counts", "backward_counts", "changed_counts"):
        if k in r:
            out[k.replace("_counts", "_states_total")] = int(sum(r[k]))
    return out
