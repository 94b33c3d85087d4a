"""Weight-truncated sparse Rank DP with meet in the middle, run on the GPU.

The call surface is that of the reference's cdist/sparse_dp.py: Trellis(h, lrow, order, nrows) / Trellis.for_query,
BACKENDS, Sweep, query(tr, K, mode, split, use_h, witness, lower, every, backend, ...), distance(p, ...),
choose_split / choose_schedule, BudgetExceeded, summary.  Its semantics (cdist/sparse_dp.py:1-30): a state of weight
above the threshold K, or whose weight plus an admissible lower bound on the faults still needed exceeds K
(lookahead, use_h), can never lie on a solution of weight <= K and is dropped; every state on such a solution keeps
its exact value, so a sweep answers "the least weight if it is <= K".  The meet in the middle splits the weight:
the backward sweep keeps weights <= K - a(c), the forward one weights <= a(c) - 1 and joins its new or improved
states (weights <= a(c)) with the backward table at every cut.

How it runs here:
  Trellis       built natively (tdp_sp_trellis: span form, slot colouring, closing masks) with the lookahead data
                from tdp_sp_bounds; a state is a key of up to 120 slots (two 64-bit words on the device).
  query         one call of libtdp's native runner (tdp_sq_run) for the "hip" backend: the whole query -- backward
                sweep, forward frontier with joins, checkpoints, traceback -- runs on one stream with device tables,
                one read-back per cut; per-cut counts come back at the end.  Any other backend object (the
                reference's backend interface, e.g. a CPU table set used as a checker) runs through Sweep below.
  Sweep         the sweep of one direction over any backend object, stepping tables one column at a time, with
                checkpoints and a traceback through recomputed stretches.
  dense=f       (the sweeps that keep whole tables, as the reference's) a table holding more than the fraction f
                of its cut's 2^width states (2^width <= dense_max) turns into a DenseTab, stepped by one gather
                kernel per column (tdp_dn_step) until it holds fewer than f / 4 of them; query(dense=...) runs its
                sweeps this way, through Sweep, on the device tables.
  split choice  probe sweeps (tdp_sq_run's profile mode) give the states of each weight at every cut; the split
                minimises a cost model of the native runner (which keeps backward tables instead of recomputing
                them when they fit its budget).
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib

INF = 255
MAX_WEIGHT = INF - 1
Q = 720720
KEYMASK = (1 << 56) - 1
MAX_SLOTS = 120                  # key bits of a device state: 64 + 56
M64 = (1 << 64) - 1
U64 = np.uint64
KEEP_BYTES = 48 << 30            # device bytes of tables a native query keeps in full before checkpointing only
DENSE_MAX_BITS = 40              # the widest dense table (tdp_dn_*)


def _lo_hi(x):
    return x & M64, x >> 64


def _join(lo, hi):
    return int(lo) | (int(hi) << 64)


# ---------------------------------------------------------------------------------------------- trellis
class Trellis:
    """The minimal-span trellis of the query min{|e| : H e = 0, (L e)_obs = 1} in a column order, built by libtdp.

    lrow: the target row over positions (bit p = the column at position p flips the observable).  Rows sit on fixed
    state bits ("slots") so that a state means the same in both directions; cut c separates positions [0, c) and
    [c, N).  Attributes (as the reference's): order, N, feasible, nslots, col[p], fsel_mask/fsel_bits[p] (rows
    closing at p going forward and the bits they must have), bsel_mask/bsel_bits[p] (going backward), tmask[c] (the
    target row's slot while active), active[c] (slots of the active rows), width[c], starts_at/ends_at[p]
    ((slot, target) of the rows starting / ending at p)."""

    def __init__(self, h, lrow, order, nrows, name=""):
        from .trellis import column_words
        self.name = name
        self.order = [int(x) for x in order]
        self.N = N = len(self.order)
        self._words = column_words(h, nrows)
        self._nrows = nrows
        self._targets = np.zeros(max(1, N), dtype=np.uint8)
        for p, j in enumerate(self.order):
            self._targets[j] = (lrow >> p) & 1
        self._ord = np.asarray(self.order or [0], dtype=np.int32)
        prm = np.zeros((max(N, 1), 10), dtype=np.uint64)
        tm = np.zeros((N + 1, 2), dtype=np.uint64)
        ns, fe = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
        vp = ctypes.c_void_p
        _lib.check(_lib.lib().tdp_sp_trellis(self._words.ctypes.data_as(vp), N, nrows, self._ord.ctypes.data_as(vp),
                                             self._targets.ctypes.data_as(vp), 1, MAX_SLOTS, prm.ctypes.data_as(vp),
                                             tm.ctypes.data_as(vp), ns.ctypes.data_as(vp), fe.ctypes.data_as(vp)),
                   "tdp_sp_trellis")
        self.feasible = bool(fe[0])
        if not self.feasible:
            return
        self.nslots = int(ns[0])
        if self.nslots > MAX_SLOTS:
            raise ValueError("trellis needs %d bits per state (at most %d supported)" % (self.nslots, MAX_SLOTS))
        self.prm, self.tm = prm, tm
        w = [[_join(prm[p, 2 * k], prm[p, 2 * k + 1]) for p in range(N)] for k in range(5)]
        self.col, self.fsel_mask, self.fsel_bits, self.bsel_mask, self.bsel_bits = w
        self.tmask = [_join(a, b) for a, b in tm]
        # a row is active at cut c when start < c <= end: it joins after its start position, leaves at its end
        act, cur = [0] * (N + 1), 0
        for p in range(N):
            cur = (cur | self.bsel_mask[p]) & ~self.fsel_mask[p]
            act[p + 1] = cur
        self.active = act
        self.width = [bin(a).count("1") for a in act]

        def pairs(masks, bits):
            return [[(s, (b >> s) & 1) for s in range(m.bit_length()) if (m >> s) & 1] for m, b in zip(masks, bits)]
        self.starts_at = pairs(self.bsel_mask, self.bsel_bits)
        self.ends_at = pairs(self.fsel_mask, self.fsel_bits)
        self._bounds = {}

    @classmethod
    def for_query(cls, p, obs, order):
        lrow = 0
        for pos, j in enumerate(order):
            lrow |= ((p.l[j] >> obs) & 1) << pos
        return cls(p.h, lrow, order, p.nrows, name="%s obs %d" % (getattr(p, "name", ""), obs))

    def words(self):
        """(prm, tm): the column words (N x 10) and target masks ((N + 1) x 2) of tdp_sq_run and tdp_spb_add."""
        return self.prm, self.tm

    def bound_arrays(self, forward):
        """The lookahead data of one direction from tdp_sp_bounds: (ncls int32[N+1], offsets int64[N+1], masks
        uint64[n, 2], weights int64[n], extra int64[N+1]), the classes of cut c at offsets[c]."""
        key = bool(forward)
        if key not in self._bounds:
            N, cap = self.N, max(1, self.nslots)
            ncls = np.zeros(N + 1, dtype=np.int32)
            extra = np.zeros(N + 1, dtype=np.int64)
            masks = np.zeros(((N + 1) * cap, 2), dtype=np.uint64)
            weights = np.zeros((N + 1) * cap, dtype=np.int64)
            vp = ctypes.c_void_p
            _lib.check(_lib.lib().tdp_sp_bounds(self._words.ctypes.data_as(vp), N, self._nrows,
                                                self._ord.ctypes.data_as(vp), self._targets.ctypes.data_as(vp),
                                                int(key), cap, ncls.ctypes.data_as(vp), extra.ctypes.data_as(vp),
                                                masks.ctypes.data_as(vp), weights.ctypes.data_as(vp)),
                       "tdp_sp_bounds")
            off = np.arange(N + 1, dtype=np.int64) * cap
            self._bounds[key] = (ncls, off, masks, weights, extra)
        return self._bounds[key]

    def bound_data(self, forward):
        """Per cut (masks, weights, extra) of the bound h(s) = ceil((extra + sum_k w_k |(s ^ tmask) & mask_k|) / Q)
        (the reference's format, cdist/sparse_dp.py:148-192)."""
        key = ("py", bool(forward))
        if key not in self._bounds:
            ncls, off, masks, weights, extra = self.bound_arrays(forward)
            out = []
            for c in range(self.N + 1):
                k0, k = int(off[c]), int(ncls[c])
                out.append((tuple(_join(*masks[i]) for i in range(k0, k0 + k)),
                            tuple(int(w) for w in weights[k0:k0 + k]), int(extra[c])))
            self._bounds[key] = out
        return self._bounds[key]


# ---------------------------------------------------------------------------------------------- device tables
class DevTable:
    """n states on the device (16 bytes each) and, once built, their hash index."""
    __slots__ = ("buf", "n", "slots", "capacity")

    def __init__(self, buf, n):
        self.buf, self.n, self.slots, self.capacity = buf, n, None, 0

    def ptr(self):
        return self.buf.data_ptr()


class DenseTab:
    """A table in dense form (the sweeps' fallback, cdist/sparse_dp.py:221-289): T[i] on the device (uint8, 255 =
    absent, 2^w entries) for the compact index i of a state, bit k of i = the state's bit at slots[k] (ascending);
    n = the states present."""
    __slots__ = ("T", "slots", "n", "mask")

    def __init__(self, T, slots, n):
        self.T, self.slots, self.n = T, list(slots), int(n)
        self.mask = sum(1 << s for s in self.slots)

    @property
    def w(self):
        return len(self.slots)

    def compact(self, s):
        """The compact index of state s (its bits outside the slots are ignored)."""
        return _compact(s, self.slots)

    def slot_array(self):
        return np.array(self.slots or [0], dtype=np.uint8)


def _compact(x, slots):
    i = 0
    for k, s in enumerate(slots):
        i |= ((x >> s) & 1) << k
    return i


def _pair_array(x):
    return (ctypes.c_uint64 * 2)(x & M64, x >> 64)


class HipTables:
    """Sparse tables on the GPU behind the reference's backend interface (cdist/sparse_dp.py:300-445, 506-605):
    start, size, nbytes, arrays, from_arrays, lookup, step, changed_size, persistent, truncate, join.  step() takes
    the lookahead data of a cut either as this backend's device form (bound()) or as the reference's host triple
    (masks, weights, extra) of Trellis.bound_data."""
    name = "hip"

    def __init__(self, tr, stream=None):
        if tr.nslots > MAX_SLOTS:
            raise ValueError("the device tables hold at most %d row slots (this trellis needs %d)"
                             % (MAX_SLOTS, tr.nslots))
        self.tr = tr
        self.L = _lib.lib()
        self.stream = stream
        self._sh = _lib.stream_handle(stream)
        self.counters = _lib.device_empty(64)
        self._host = np.zeros(8, dtype=np.uint64)
        self._dev_bounds = {}

    def _read(self, k):
        _lib.check(self.L.tdp_read_back(self._host.ctypes.data_as(ctypes.c_void_p), self.counters.data_ptr(), 8 * k,
                                        self._sh), "tdp_read_back")
        return self._host[:k].tolist()

    @staticmethod
    def _buffer(nstates):
        return _lib.device_empty(16 * max(1, nstates))

    def index(self, tab):
        if tab.slots is None:
            cap = 64
            while cap < 2 * tab.n:
                cap *= 2
            tab.slots, tab.capacity = _lib.device_empty(4 * cap), cap
            _lib.check(self.L.tdp_sp_index(tab.ptr(), tab.n, tab.slots.data_ptr(), cap, self._sh), "tdp_sp_index")
        return tab

    def start(self):
        return self.from_arrays([0], np.zeros(1, dtype=np.uint8))

    @staticmethod
    def size(tab):
        return tab.n

    @staticmethod
    def nbytes(tab):
        return (1 << tab.w) if isinstance(tab, DenseTab) else 16 * tab.n

    def arrays(self, tab):
        """(keys, vals): keys as uint64 when they all fit 64 bits, else Python ints (object array)."""
        if isinstance(tab, DenseTab):
            tab = self.from_dense(tab)
        if not tab.n:
            return np.zeros(0, dtype=U64), np.zeros(0, dtype=np.uint8)
        raw = tab.buf[:16 * tab.n].cpu().numpy().view(U64).reshape(-1, 2)
        vals = (raw[:, 1] >> U64(56)).astype(np.uint8)
        hi = raw[:, 1] & U64(KEYMASK)
        if hi.any():
            return np.array([_join(a, b) for a, b in zip(raw[:, 0], hi)], dtype=object), vals
        return raw[:, 0].copy(), vals

    def from_arrays(self, keys, vals):
        n = len(vals)
        raw = np.zeros((max(1, n), 2), dtype=U64)
        for i, (k, v) in enumerate(zip(keys, vals)):
            lo, hi = _lo_hi(int(k))
            raw[i] = (lo, hi | (int(v) << 56))
        buf = self._buffer(n)
        buf.copy_(_lib.torch().from_numpy(raw.view(np.uint8).reshape(-1)))
        return DevTable(buf, n)

    def lookup_many(self, tab, keys):
        m = len(keys)
        if not m or not tab.n:
            return np.full(m, INF, dtype=np.uint8)
        if isinstance(tab, DenseTab):
            inside = [not (int(k) & ~tab.mask) for k in keys]
            idx = [tab.compact(int(k)) for k, ok in zip(keys, inside) if ok]
            out = np.full(m, INF, dtype=np.uint8)
            if idx:
                t = _lib.torch()
                got = tab.T[t.tensor(idx, dtype=t.int64, device=tab.T.device)].cpu().numpy()
                out[np.flatnonzero(inside)] = got
            return out
        self.index(tab)
        q = np.array([_lo_hi(int(k)) for k in keys], dtype=U64)
        dq = _lib.torch().from_numpy(q.view(np.uint8).reshape(-1)).cuda()
        dv = _lib.device_empty(m)
        _lib.check(self.L.tdp_sp_lookup(tab.ptr(), tab.slots.data_ptr(), tab.capacity, dq.data_ptr(), m,
                                        dv.data_ptr(), self._sh), "tdp_sp_lookup")
        return dv[:m].cpu().numpy()

    def lookup(self, tab, s):
        return int(self.lookup_many(tab, [s])[0])

    def _device_bound(self, hd):
        """The device form (mask tensor, weight tensor, ncls, extra) of one cut's (masks, weights, extra)."""
        key = id(hd)
        got = self._dev_bounds.get(key)
        if got is None or got[0] is not hd:
            masks, weights, extra = hd
            t = _lib.torch()
            m = np.array([_lo_hi(int(x)) for x in masks] or [(0, 0)], dtype=U64).reshape(-1)
            w = np.array([int(x) for x in weights] or [0], dtype=np.int64)
            got = (hd, (t.from_numpy(m).cuda(), t.from_numpy(w).cuda(), len(masks), int(extra)))
            self._dev_bounds[key] = got
        return got[1]

    def bound(self, forward):
        """Per cut, the lookahead data in this backend's device form."""
        return [self._device_bound(hd) for hd in self.tr.bound_data(forward)]

    def step(self, tab, upd, sel_mask, sel_bits, cap, K, tmask, hd, keep_below=None, want_changed=False):
        """T'[s] = min(T[s], T[s ^ upd] + 1), then the closing-row filter, the cap, keep_below and the lookahead:
        (table, changed) as cdist/sparse_dp.py:338-403 returns them."""
        self.index(tab)
        n = tab.n
        out = self._buffer(2 * n)
        ch = self._buffer(2 * n) if want_changed else None
        if hd is None:
            ncls, mp, wp, extra, use_h = 0, None, None, 0, 0
        else:
            dm, dw, ncls, extra = hd if len(hd) == 4 else self._device_bound(hd)
            mp, wp, use_h = dm.data_ptr(), dw.data_ptr(), 1
        _lib.check(self.L.tdp_sp_step(tab.ptr(), n, tab.slots.data_ptr(), tab.capacity, _pair_array(upd),
                                      _pair_array(sel_mask), _pair_array(sel_bits), cap, keep_below or 0, ncls, mp,
                                      wp, extra, _pair_array(tmask), K, use_h, out.data_ptr(), 2 * n,
                                      ch.data_ptr() if ch is not None else None, 2 * n, self.counters.data_ptr(),
                                      self._sh), "tdp_sp_step")
        nout, nch, overflow = self._read(3)
        if overflow:
            raise RuntimeError("sparse step output overflow")
        return DevTable(out, nout), (DevTable(ch, nch) if want_changed else None)

    @staticmethod
    def changed_size(ch):
        return ch.n

    def persistent(self, tab):
        """The table in a buffer of its own size (step outputs are sized for twice their input; a dense step writes
        a new table, so a dense one is kept as it is)."""
        if isinstance(tab, DenseTab):
            return tab
        if tab.buf.numel() > 16 * max(1, tab.n) * 5 // 4 + 16:
            buf = self._buffer(tab.n)
            buf[:16 * tab.n].copy_(tab.buf[:16 * tab.n])
            return DevTable(buf, tab.n)
        return tab

    def truncate(self, tab, below):
        out = self._buffer(tab.n)
        _lib.check(self.L.tdp_sp_truncate(tab.ptr(), tab.n, below, out.data_ptr(), self.counters.data_ptr(),
                                          self._sh), "tdp_sp_truncate")
        return DevTable(out, self._read(1)[0])

    def join(self, ch, btab, tmask):
        """(total, forward key, forward value, backward key) of the least value(f) + value(b) over changed states f
        and backward states b = f ^ tmask (ties: the least forward key), or None."""
        if not ch.n or not btab.n:
            return None
        if isinstance(btab, DenseTab):
            sl = btab.slot_array()
            _lib.check(self.L.tdp_dn_join(ch.ptr(), ch.n, btab.T.data_ptr(), btab.w, sl.ctypes.data,
                                          _pair_array(tmask), self.counters.data_ptr(), self._sh), "tdp_dn_join")
            return self._join_result(tmask)
        self.index(btab)
        _lib.check(self.L.tdp_sp_join(ch.ptr(), ch.n, btab.ptr(), btab.slots.data_ptr(), btab.capacity,
                                      _pair_array(tmask), self.counters.data_ptr(), self._sh), "tdp_sp_join")
        return self._join_result(tmask)

    def _join_result(self, tmask):
        tot, hi, lo, fval = self._read(4)
        if tot == M64:
            return None
        fkey = _join(lo, hi)
        return int(tot), fkey, int(fval), fkey ^ tmask

    # ---- dense form (cdist/sparse_dp.py:221-289, 680-707)
    @staticmethod
    def is_dense(tab):
        return isinstance(tab, DenseTab)

    def to_dense(self, tab, slots):
        """The sparse table `tab` as a DenseTab over the ascending `slots` (its states' bits lie in them)."""
        dt = DenseTab(_lib.device_empty(max(4, 1 << len(slots))), slots, tab.n)
        sl = dt.slot_array()                          # held until the call returns
        _lib.check(self.L.tdp_dn_from_sparse(tab.ptr(), tab.n, sl.ctypes.data, dt.w, dt.T.data_ptr(),
                                             self._sh), "tdp_dn_from_sparse")
        return dt

    def from_dense(self, dt):
        """The DenseTab's states as a sparse table."""
        out = self._buffer(dt.n)
        sl = dt.slot_array()
        _lib.check(self.L.tdp_dn_to_sparse(dt.T.data_ptr(), dt.w, sl.ctypes.data, out.data_ptr(),
                                           max(1, dt.n), self.counters.data_ptr(), self._sh), "tdp_dn_to_sparse")
        n, over = self._read(2)
        if over or n != dt.n:
            raise RuntimeError("dense table holds %d states, %d expected" % (n, dt.n))
        return DevTable(out, n)

    def dense_step(self, dt, starts, ends, upd, cap, K, tmask, hd):
        """One column step of a dense table: rows `starts` ((slot, target) pairs) get new bits, absent where set;
        T'[x] = min(T[x], T[x ^ upd] + 1); rows `ends` are fixed at their target and removed; states above cap or,
        with hd = (masks, weights, extra), whose value plus the lookahead bound exceeds K are dropped."""
        old = dt.slots
        new_rows = sorted(s for s, _ in starts)
        if set(new_rows) & set(old):
            raise AssertionError("a starting row's slot is already active")
        mid = sorted(old + new_rows)
        pos = {s: k for k, s in enumerate(mid)}
        ends = sorted(ends)
        if any(s not in pos for s, _ in ends):
            raise AssertionError("an ending row's slot is not active")
        gone = {s for s, _ in ends}
        new = [s for s in mid if s not in gone]
        start_pos = np.array(sorted((pos[s] for s in new_rows), reverse=True) or [0], dtype=np.uint8)
        end_pos = np.array([pos[s] for s, _ in ends] or [0], dtype=np.uint8)
        end_bit = np.array([t for _, t in ends] or [0], dtype=np.uint8)
        ncls, masks, weights, extra, ct = 0, np.zeros(1, dtype=U64), np.zeros(1, dtype=np.int64), 0, 0
        if hd is not None:
            by_mask = {}
            for m, wt in zip(hd[0], hd[1]):
                cm = _compact(int(m), new)
                if cm:
                    by_mask[cm] = by_mask.get(cm, 0) + int(wt)
            ncls, extra, ct = len(by_mask), int(hd[2]), _compact(int(tmask), new)
            if by_mask:
                masks = np.array(list(by_mask), dtype=U64)
                weights = np.array(list(by_mask.values()), dtype=np.int64)
        U = _lib.device_empty(max(4, 1 << len(new)))
        self._dense_launch(dt.T.data_ptr(), dt.w, U.data_ptr(), len(mid), start_pos.ctypes.data, len(new_rows),
                           end_pos.ctypes.data, end_bit.ctypes.data, len(ends), _compact(int(upd), mid), cap, K, ncls,
                           masks.ctypes.data, weights.ctypes.data, extra, ct, int(hd is not None))
        return DenseTab(U, new, self._read(1)[0])

    def _dense_launch(self, *args):
        """One tdp_dn_step call on this backend's stream (args up to use_h; a method of its own so that a caller
        can time the launch alone)."""
        _lib.check(self.L.tdp_dn_step(*args, self.counters.data_ptr(), self._sh), "tdp_dn_step")

    def hist(self, tab):
        h = _lib.device_empty(8 * 256)
        _lib.check(self.L.tdp_sp_hist(tab.ptr(), tab.n, h.data_ptr(), self._sh), "tdp_sp_hist")
        return h.cpu().numpy().view(U64).astype(np.int64)


BACKENDS = {"hip": HipTables, "c": HipTables}
NATIVE = ("hip", "c")


class BudgetExceeded(RuntimeError):
    """A table outgrew max_states (kind "states") or a sweep ran past its deadline (kind "deadline")."""

    def __init__(self, message, kind):
        super().__init__(message)
        self.kind = kind


def _check_caps(caps, direction):
    if max(caps) > MAX_WEIGHT:
        raise ValueError("a sweep keeping weights up to %d: the tables hold at most %d (8-bit values)"
                         % (max(caps), MAX_WEIGHT))
    pairs = zip(caps, caps[1:])
    if not all((a >= b) if direction > 0 else (a <= b) for a, b in pairs):
        raise AssertionError("forward caps must not increase" if direction > 0 else "backward caps must not decrease")


# ---------------------------------------------------------------------------------------------- one sweep
class Sweep:
    """One direction of the truncated sweep over a table backend (any object with the reference's backend
    interface), with tables kept at checkpoints.

    direction +1 runs cuts 0 -> N from {0: 0}, -1 runs N -> 0; caps[c] (or one cap) bounds the weights kept at cut c,
    non-increasing in the sweep's direction.  frontier (the forward half of a meet in the middle): the step out of
    cut p keeps weights <= caps[p] - 1 and reports the new or improved states of weight <= caps[p] as `changed`;
    where the caps drop, the table before the drop is reported as `extra`.  Attributes after run(): counts,
    changed_counts, peak, peak_bytes, steps_done, ckpt (cut -> kept table)."""

    def __init__(self, tr, direction, K, cap=None, use_h=True, every=None, frontier=False, backend="hip",
                 dense=None, dense_max=1 << 30, max_states=None, deadline=None, caps=None):
        self.tr, self.direction, self.K = tr, direction, K
        self.caps = [min(int(x), K) for x in caps] if caps is not None else [min(cap, K)] * (tr.N + 1)
        _check_caps(self.caps, direction)
        self.use_h, self.frontier = use_h, frontier
        self.be = BACKENDS[backend](tr) if isinstance(backend, str) else backend
        # the dense form is for the sweeps that keep whole tables (the frontier's changed states stay sparse)
        self.dense = dense if dense and not frontier else None
        self.dense_max = min(int(dense_max), 1 << DENSE_MAX_BITS)
        if self.dense is not None and not hasattr(self.be, "to_dense"):
            raise ValueError("backend %r has no dense form" % getattr(self.be, "name", self.be))
        self.hd = self._bounds() if use_h else None
        self.every = every or max(1, math.isqrt(tr.N))
        self.max_states, self.deadline = max_states, deadline
        self.ckpt = {}
        self.counts = [0] * (tr.N + 1)
        self.changed_counts = [0] * (tr.N + 1)
        self.dense_cuts = self.peak = self.peak_bytes = self.steps_done = 0

    def _bounds(self):
        fwd = self.direction > 0
        if hasattr(self.be, "bound"):
            return self.be.bound(fwd)
        if getattr(self.be, "name", "") == "c" and hasattr(self.tr, "bound_arrays"):
            ncls, off, masks, weights, extra = self.tr.bound_arrays(fwd)
            return [(masks[o:o + k, 0].copy(), weights[o:o + k].copy(), int(e))
                    for k, o, e in zip(ncls.tolist(), off.tolist(), extra.tolist())]
        return self.tr.bound_data(fwd)

    def first(self):
        return (0 if self.direction > 0 else self.tr.N), self.be.start()

    def _advance(self, tab, p, want_changed):
        """The column step at position p: (new cut, table, changed, extra)."""
        tr = self.tr
        if self.direction > 0:
            cut, sel = p + 1, (tr.fsel_mask[p], tr.fsel_bits[p])
        else:
            cut, sel = p, (tr.bsel_mask[p], tr.bsel_bits[p])
        self.steps_done += 1
        cap = self.caps[p] if self.frontier else self.caps[cut]
        dense_in = self.dense is not None and self.be.is_dense(tab)
        if dense_in and (1 << tr.width[cut]) > self.dense_max:
            tab, dense_in = self.be.from_dense(tab), False          # too wide to stay dense
        if dense_in:
            fwd = self.direction > 0
            hd = tr.bound_data(fwd)[cut] if self.use_h else None
            tab = self.be.dense_step(tab, tr.starts_at[p] if fwd else tr.ends_at[p],
                                     tr.ends_at[p] if fwd else tr.starts_at[p], tr.col[p], cap, self.K,
                                     tr.tmask[cut], hd)
            if self.be.size(tab) < self.dense / 4 * (1 << len(tab.slots)):
                tab = self.be.from_dense(tab)
            return cut, tab, None, None
        keep = cap if self.frontier else None
        tab, changed = self.be.step(tab, tr.col[p], sel[0], sel[1], cap, self.K, tr.tmask[cut],
                                    self.hd[cut] if self.hd else None, keep, want_changed)
        if self.frontier and self.caps[cut] < cap:
            return cut, self.be.truncate(tab, self.caps[cut]), changed, tab
        if self.dense is not None:
            full = 1 << tr.width[cut]
            if self.be.size(tab) > self.dense * full and full <= self.dense_max:
                tab = self.be.to_dense(tab, [s for s in range(tr.active[cut].bit_length())
                                             if (tr.active[cut] >> s) & 1])
        return cut, tab, changed, None

    def _steps(self, cut, tab, stop, want_changed=False):
        """(cut, table, changed, extra) for each cut after `cut` up to `stop`."""
        positions = range(cut, stop) if self.direction > 0 else range(cut - 1, stop - 1, -1)
        for p in positions:
            cut, tab, changed, extra = self._advance(tab, p, want_changed)
            yield cut, tab, changed, extra

    def is_ckpt(self, c):
        return c == 0 or c == self.tr.N or c % self.every == 0

    def run(self, on_cut=None):
        """The whole sweep, recording counts and checkpoints; on_cut(cut, table, changed, extra) may return True to
        stop.  Returns the last cut reached."""
        c, tab = self.first()
        self.ckpt[c] = tab
        self.counts[c] = self.changed_counts[c] = self.peak = 1
        kept_bytes = self.be.nbytes(tab)
        if on_cut and on_cut(c, tab, tab, None):
            return c
        end = self.tr.N if self.direction > 0 else 0
        for c, tab, changed, extra in self._steps(c, tab, end, on_cut is not None):
            n = self.be.size(tab)
            nc = self.be.changed_size(changed) if changed is not None else 0
            self.counts[c], self.changed_counts[c] = n, nc
            self.peak = max(self.peak, n, nc)
            if self.dense is not None and self.be.is_dense(tab):
                self.dense_cuts += 1
            if self.is_ckpt(c):
                self.ckpt[c] = self.be.persistent(tab)
                kept_bytes += self.be.nbytes(tab)
            self.peak_bytes = max(self.peak_bytes, kept_bytes + self.be.nbytes(tab))
            if self.max_states is not None and max(n, nc) > self.max_states:
                raise BudgetExceeded("%d states at cut %d" % (max(n, nc), c), "states")
            if self.deadline is not None and time.monotonic() > self.deadline:
                raise BudgetExceeded("deadline at cut %d" % c, "deadline")
            if on_cut and on_cut(c, tab, changed, extra):
                return c
        return end

    def segment(self, lo, hi):
        """Tables at cuts lo..hi, recomputed from the checkpoint that precedes them in the sweep."""
        fwd = self.direction > 0
        base = max(c for c in self.ckpt if c <= lo) if fwd else min(c for c in self.ckpt if c >= hi)
        tab = self.ckpt[base]
        out = {base: tab} if lo <= base <= hi else {}
        for c, t, _, _ in self._steps(base, tab, hi if fwd else lo):
            if lo <= c <= hi:
                out[c] = self.be.persistent(t)
        return out

    def table(self, c, cache):
        """The table at cut c for requests running against the sweep (cache: a dict the caller keeps): a whole
        stretch between checkpoints is recomputed at once."""
        if c not in cache:
            cache.clear()
            if self.direction > 0:
                cache.update(self.segment(max(x for x in self.ckpt if x <= c), c))
            else:
                cache.update(self.segment(c, min(x for x in self.ckpt if x >= c)))
        return cache[c]

    def traceback(self, cut, state, value):
        """Columns of a partial solution reaching `state` with weight `value` at `cut`: walking back to the sweep's
        start, a column was used where the table before its step does not hold the state with the same value."""
        tr, cache, used = self.tr, {}, []
        s, v = int(state), int(value)
        fwd = self.direction > 0
        for p in (range(cut - 1, -1, -1) if fwd else range(cut, tr.N)):
            tab = self.table(p if fwd else p + 1, cache)
            before = s | (tr.fsel_bits[p] if fwd else tr.bsel_bits[p])
            if self.be.lookup(tab, before) == v:
                s = before
                continue
            if not tr.col[p] or self.be.lookup(tab, before ^ tr.col[p]) != v - 1:
                raise AssertionError("traceback lost the path at position %d" % p)
            s, v = before ^ tr.col[p], v - 1
            used.append(tr.order[p])
        if (s, v) != (0, 0):
            raise AssertionError("traceback ended at state %x value %d" % (s, v))
        return used


# ---------------------------------------------------------------------------------------------- native runs
class _Spec(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("N", "K", "mode", "use_h", "witness", "lower", "every", "hist_top")] + [
        ("max_states", ctypes.c_uint64), ("max_seconds", ctypes.c_double), ("keep_bytes", ctypes.c_uint64),
        ("prm", ctypes.c_void_p), ("tm", ctypes.c_void_p), ("split", ctypes.c_void_p)] + [
        (k, ctypes.c_void_p * 2) for k in ("bnd_ncls", "bnd_off", "bnd_mask", "bnd_w", "bnd_extra")] + [
        ("hist", ctypes.c_void_p), ("dev_alloc", _lib.ALLOC_FN), ("dev_free", _lib.FREE_FN),
        ("alloc_ctx", ctypes.c_void_p)]


class _Result(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("status", "fail_cut", "value", "join_cut", "nwitness", "pad")] + [
        (k, ctypes.c_uint64) for k in ("steps", "joined", "peak_states", "peak_bytes")] + [
        (k, ctypes.c_double) for k in ("seconds", "backward_seconds", "forward_seconds", "traceback_seconds")]


SQ_FORWARD, SQ_MITM, SQ_PROFILE_FW, SQ_PROFILE_BW = range(4)
SQ_DONE, SQ_STATES, SQ_DEADLINE = range(3)


def _native(tr, K, mode, split=None, use_h=False, witness=False, lower=0, every=None, max_states=None,
            deadline=None, hist_top=None, keep_bytes=None, stream=None):
    """One tdp_sq_run call: (result record, forward counts, backward counts, changed counts, witness positions,
    histogram or None)."""
    N = tr.N
    spec = _Spec()
    spec.N, spec.K, spec.mode = N, K, mode
    spec.use_h, spec.witness, spec.lower = int(bool(use_h)), int(bool(witness)), int(lower)
    spec.every = int(every or max(1, math.isqrt(N)))
    spec.max_states = int(max_states or 0)
    spec.max_seconds = -1.0 if deadline is None else max(0.0, deadline - time.monotonic())
    spec.keep_bytes = KEEP_BYTES if keep_bytes is None else int(keep_bytes)
    prm, tm = tr.words()
    prm, tm = np.ascontiguousarray(prm), np.ascontiguousarray(tm)
    spec.prm, spec.tm = prm.ctypes.data, tm.ctypes.data
    keep = [prm, tm]
    if split is not None:
        sp = np.ascontiguousarray(split, dtype=np.int32)
        spec.split = sp.ctypes.data
        keep.append(sp)
    if use_h:
        for d, fwd in enumerate((True, False)):
            arrays = tr.bound_arrays(fwd)
            keep += arrays
            ncls, off, masks, weights, extra = arrays
            spec.bnd_ncls[d], spec.bnd_off[d] = ncls.ctypes.data, off.ctypes.data
            spec.bnd_mask[d], spec.bnd_w[d], spec.bnd_extra[d] = masks.ctypes.data, weights.ctypes.data, extra.ctypes.data
    hist = None
    if hist_top is not None:
        spec.hist_top = hist_top
        hist = np.zeros((N + 1, hist_top + 1), dtype=np.uint64)
        spec.hist = hist.ctypes.data
    fw = np.zeros(N + 1, dtype=np.uint64)
    bw = np.zeros(N + 1, dtype=np.uint64)
    ch = np.zeros(N + 1, dtype=np.uint64)
    wit = np.zeros(K + 2, dtype=np.int32)
    res = _Result()
    # the query's tables come from torch's caching allocator: returning tens of GB to the driver after every query
    # and mapping them again in the next cost seconds
    with _lib.TorchDeviceAllocator(stream) as arena:
        spec.dev_alloc, spec.dev_free = arena.alloc_fn, arena.free_fn
        rc = _lib.lib().tdp_sq_run(ctypes.byref(spec), fw.ctypes.data, bw.ctypes.data, ch.ctypes.data,
                                   wit.ctypes.data, ctypes.byref(res), _lib.stream_handle(stream))
    _lib.check(rc, "tdp_sq_run")
    if res.status == SQ_STATES:
        raise BudgetExceeded("%d states at cut %d" % (res.peak_states, res.fail_cut), "states")
    if res.status == SQ_DEADLINE:
        raise BudgetExceeded("deadline at cut %d" % res.fail_cut, "deadline")
    return res, fw.tolist(), bw.tolist(), ch.tolist(), wit[:res.nwitness].tolist(), hist


# ---------------------------------------------------------------------------------------------- split choice
def cut_profiles(tr, direction, cap, backend="hip"):
    """States of each weight <= cap at every cut, array (N + 1, cap + 1), of a sweep keeping weights <= cap
    without lookahead."""
    if isinstance(backend, str) and backend in NATIVE:
        *_, hist = _native(tr, cap, SQ_PROFILE_FW if direction > 0 else SQ_PROFILE_BW, hist_top=cap)
        return hist.astype(np.float64)
    sw = Sweep(tr, direction, cap, cap, use_h=False, backend=backend)
    c0, tab = sw.first()
    out = np.zeros((tr.N + 1, cap + 1))
    out[c0, 0] = 1
    for c, tab, _, _ in sw._steps(c0, tab, tr.N if direction > 0 else 0):
        vals = np.asarray(sw.be.arrays(tab)[1], dtype=np.int64)
        out[c] = np.bincount(vals, minlength=cap + 1)[:cap + 1]
    return out


def weight_profile(tr, direction, cap, backend="hip"):
    """States of each weight <= cap summed over the cuts of such a sweep."""
    return cut_profiles(tr, direction, cap, backend).sum(axis=0)


def probe_profiles(tr, backend="hip", probe=3, budget=2e9):
    """Per-cut state counts by weight in both directions, measured up to weight `probe`, or up to 2 in a direction
    where weight `probe` would cost more than `budget` states (estimated from the growth between weights 1 and 2;
    budget may be a number or {+1: ..., -1: ...}).  Cached on the trellis; returns ({+1: array, -1: array},
    seconds spent now)."""
    t0 = time.time()
    cache = tr.__dict__.setdefault("_profiles", {})
    bname = backend if isinstance(backend, str) else getattr(backend, "name", "backend")
    out = {}
    for d in (+1, -1):
        lim = budget[d] if isinstance(budget, dict) else budget
        low = cache.get((d, 2, bname))
        if low is None:
            low = cache[(d, 2, bname)] = cut_profiles(tr, d, min(2, probe), backend)
        tot = low.sum(axis=0)
        grown = tot[-1] * (tot[-1] / tot[1]) ** (probe - 2) if len(tot) > 2 and tot[1] > 0 else np.inf
        if probe > 2 and grown < lim:
            key = (d, probe, bname)
            if key not in cache:
                cache[key] = cut_profiles(tr, d, probe, backend)
            out[d] = cache[key]
        else:
            out[d] = low
    return out, time.time() - t0


def _extend(per_weight, K, window=0):
    """Counts by weight extended to weights 0..K: each new weight is the last one times the growth between the two
    measured ones (at least 1), smoothed over +-window cuts when per_weight is per cut."""
    p = np.atleast_2d(per_weight).astype(np.float64)
    top = p.shape[1] - 1
    if top >= K:
        return p[:, :K + 1]
    last, prev = p[:, top], p[:, top - 1] if top >= 1 else p[:, top]
    if window:
        def windowed(x):                 # sums over cuts c - window .. c + window
            cs = np.concatenate([[0.0], np.cumsum(x)])
            i = np.arange(x.size)
            return cs[np.minimum(x.size, i + window + 1)] - cs[np.maximum(0, i - window)]
        last, prev = windowed(last), windowed(prev)
    growth = np.maximum(1.0, np.divide(last, prev, out=np.ones_like(last), where=prev > 0))
    cols = [p[:, w] for w in range(top + 1)]
    while len(cols) <= K:
        cols.append(cols[-1] * growth)
    return np.stack(cols, axis=1)


def _native_cost(F, B, K, a, backward_factor):
    """Predicted states of the native meet in the middle at split a, from states by weight (forward F, backward B,
    per cut or summed): the forward tables (weights <= a - 1) and frontier (weight a), the backward tables (weights
    <= K - a), the latter twice when they must be recomputed from checkpoints."""
    cF, cB = np.cumsum(F, axis=-1), np.cumsum(B, axis=-1)
    return cF[..., a - 1] + F[..., a - 1] + backward_factor * cB[..., K - a]


def choose_split(tr, K, backend="hip", probe=3, budget=2e9):
    """A uniform split a for threshold K minimising the predicted states of the native runner (returns (a, info)).
    The backward tables count once when the sweep is predicted to keep them all (KEEP_BYTES), twice otherwise."""
    prof, secs = probe_profiles(tr, backend, probe, budget)
    F, B = _extend(prof[+1].sum(axis=0), K)[0], _extend(prof[-1].sum(axis=0), K)[0]
    costs = {}
    for a in range(1, K + 1):
        factor = 1.0 if 48 * np.cumsum(B)[K - a] <= KEEP_BYTES else 2.0
        costs[a] = float(_native_cost(F, B, K, a, factor))
    a = min(costs, key=costs.get)
    return a, {"forward_profile": prof[+1].sum(axis=0).tolist(), "backward_profile": prof[-1].sum(axis=0).tolist(),
               "predicted_work": costs, "probe_seconds": secs}


def choose_schedule(tr, K, backend="hip", probe=3, budget=2e9, window=100):
    """A split a(c) per cut, non-increasing in c, minimising the summed per-cut predicted states (the per-cut form
    of choose_split's model, with the probe counts extended over a window of cuts).  The search runs from the last
    cut back: best[c][a] = cost_c(a) + min over a' <= a of best[c + 1][a'], then a(0) = argmin best[0] and each
    a(c + 1) = the argmin of best[c + 1] over a' <= a(c).  Returns (schedule, info) with the predicted work, peak and
    mean states and device bytes."""
    prof, secs = probe_profiles(tr, backend, probe, budget)
    F, B = _extend(prof[+1], K, window), _extend(prof[-1], K, window)
    n = tr.N + 1
    factor = 1.0 if 48 * np.cumsum(B, axis=1)[:, K - 1].sum() <= KEEP_BYTES else 2.0
    cost = np.full((n, K + 1), np.inf)
    for a in range(1, K + 1):
        cost[:, a] = _native_cost(F, B, K, a, factor)
    best = np.empty((n, K + 1))
    best[n - 1] = cost[n - 1]
    for c in range(n - 2, -1, -1):
        best[c] = cost[c] + np.minimum.accumulate(best[c + 1])
    sched = [int(np.argmin(best[0]))]
    for c in range(1, n):
        sched.append(int(np.argmin(best[c][:sched[-1] + 1])))
    av, idx = np.array(sched), np.arange(n)
    fwd = np.cumsum(F, axis=1)[idx, av - 1] + F[idx, av - 1]
    bwd = np.cumsum(B, axis=1)[idx, K - av]
    uniform = {a: float(cost[:, a].sum()) for a in range(1, K + 1)}
    return sched, {"predicted_work": float(best[0].min()), "uniform_predicted_work": uniform,
                   "predicted_peak_states": float(max(fwd.max(), bwd.max())),
                   "predicted_mean_states": float((fwd.sum() + bwd.sum()) / (2 * n)),
                   "predicted_bytes": float(48 * (fwd.sum() + bwd.sum())),
                   "schedule": _runs(sched), "forward_profile": prof[+1].sum(axis=0).tolist(),
                   "backward_profile": prof[-1].sum(axis=0).tolist(), "probe_seconds": secs}


def _runs(x):
    """A per-cut list as 'value' or 'value@cut,value@cut,...' at the cuts where it changes."""
    marks = [(v, c) for c, v in enumerate(x) if c == 0 or v != x[c - 1]]
    return str(marks[0][0]) if len(marks) == 1 else ",".join("%d@%d" % m for m in marks)


def _split_schedule(tr, K, split, backend):
    """(per-cut split, split_choice info or None, predicted work or None)."""
    info = predicted = None
    if split == "auto":
        split, info = choose_split(tr, K, backend)
        predicted = info["predicted_work"][split]
    elif split == "schedule":
        split, info = choose_schedule(tr, K, backend)
        predicted = info["predicted_work"]
    if split is None:
        split = K // 2 + 1
    if isinstance(split, (list, tuple, np.ndarray)):
        sched = [max(1, min(int(x), K)) for x in split]
    else:
        sched = [max(1, min(int(split), K))] * (tr.N + 1)
    if len(sched) != tr.N + 1 or any(a < b for a, b in zip(sched, sched[1:])):
        raise AssertionError("split must not increase")
    return sched, info, predicted


# ---------------------------------------------------------------------------------------------- one query
def query(tr, K, mode="mitm", split=None, use_h=True, witness=True, lower=0, every=None, backend="hip",
          dense=None, dense_max=1 << 30, log=None, max_states=None, deadline=None):
    """The least weight of a solution of the trellis' query if it is <= K (value None otherwise), with per-cut
    statistics and a witness (problem columns); the record of cdist/sparse_dp.py:985-1095.

    mode "forward": one sweep keeping weights <= K.  mode "mitm": split a(c) per cut (an int, "auto", "schedule"
    or N + 1 non-increasing values; default K // 2 + 1).  lower: a known lower bound, reaching it stops the sweep.
    max_states / deadline (time.monotonic()) raise BudgetExceeded.  Tables hold weights <= 254: a larger cap raises
    ValueError.  Why every solution of weight <= K is seen is argued in the reference's docstring: both bounds
    K - a(c) and a(c) fall along the cuts while a solution's left weight rises by steps of one."""
    if mode not in ("forward", "mitm"):
        raise ValueError("mode must be 'forward' or 'mitm'")
    t0 = time.time()
    bname = backend if isinstance(backend, str) else backend.name
    rec = {"K": K, "mode": mode, "use_h": bool(use_h), "value": None, "witness": None, "width": max(tr.width),
           "backend": bname}
    if mode == "forward":
        _check_caps([K], +1)
        sched = None
    else:
        sched, info, predicted = _split_schedule(tr, K, split, backend)
        if info is not None:
            rec["split_choice"] = info
        _check_caps(sched, +1)
        _check_caps([K - a for a in sched], -1)
        if log:
            log("K=%d split a=%s: backward sweep (weights <= %s)%s" % (
                K, _runs(sched), _runs([K - a for a in sched]),
                "" if predicted is None else ", predicted work %.3g" % predicted))
    if isinstance(backend, str) and backend in NATIVE and dense is None:
        return _native_query(tr, K, mode, sched, use_h, witness, lower, every, max_states, deadline, rec, t0)
    return _sweep_query(tr, K, mode, sched, use_h, witness, lower, every, backend, max_states, deadline, rec, t0,
                        log, dense, dense_max)


def _native_query(tr, K, mode, sched, use_h, witness, lower, every, max_states, deadline, rec, t0):
    res, fw, bw, ch, pos, _ = _native(tr, K, SQ_FORWARD if mode == "forward" else SQ_MITM, sched, use_h, witness,
                                      lower, every, max_states, deadline)
    common = {"peak_states": int(res.peak_states), "peak_table_bytes": int(res.peak_bytes), "steps": int(res.steps),
              "dense_cuts": 0}
    if mode == "forward":
        rec.update(forward_counts=fw, sweep_seconds=res.forward_seconds, **common)
    else:
        rec.update(forward_counts=fw, backward_counts=bw, changed_counts=ch, backward_seconds=res.backward_seconds,
                   forward_seconds=res.forward_seconds, split=sched[0] if len(set(sched)) == 1 else _runs(sched),
                   split_values=sched, joined=int(res.joined), **common)
    if res.value >= 0:
        rec["value"] = int(res.value)
        if mode == "mitm":
            rec["join_cut"] = int(res.join_cut)
        if witness:
            rec["witness"] = sorted(tr.order[p] for p in pos)
            rec["traceback_seconds"] = res.traceback_seconds
    rec["seconds"] = time.time() - t0
    return rec


def _sweep_query(tr, K, mode, sched, use_h, witness, lower, every, backend, max_states, deadline, rec, t0, log,
                 dense=None, dense_max=1 << 30):
    """query() through the sweeps of this module, one column step per backend call: for a backend object, or for
    the device tables with the dense fallback (dense=f; the forward sweep, or a meet in the middle's backward one)."""
    def fresh():
        return BACKENDS[backend](tr) if isinstance(backend, str) else backend.__class__(tr)
    if mode == "forward":
        fw = Sweep(tr, +1, K, K, use_h, every, backend=fresh(), max_states=max_states, deadline=deadline, dense=dense,
                   dense_max=dense_max)
        fw.run()
        rec.update(forward_counts=fw.counts, peak_states=fw.peak, peak_table_bytes=fw.peak_bytes,
                   sweep_seconds=time.time() - t0, steps=fw.steps_done, dense_cuts=fw.dense_cuts)
        v = fw.be.lookup(fw.ckpt[tr.N], 0)
        if v != INF:
            rec["value"] = v
            if witness:
                t1 = time.time()
                rec["witness"] = sorted(fw.traceback(tr.N, 0, v))
                rec["traceback_seconds"] = time.time() - t1
        rec["seconds"] = time.time() - t0
        return rec
    bw = Sweep(tr, -1, K, use_h=use_h, every=every, backend=fresh(), max_states=max_states, deadline=deadline,
               caps=[K - a for a in sched], dense=dense, dense_max=dense_max)
    fw = Sweep(tr, +1, K, use_h=use_h, every=bw.every, frontier=True, backend=fresh(), max_states=max_states,
               deadline=deadline, caps=sched)
    bw.run()
    t1 = time.time()
    if log:
        log("backward sweep done: %.1f s, peak %d states" % (t1 - t0, bw.peak))
    state = {"best": None, "joined": 0}
    cache = {}

    def at_cut(c, tab, changed, extra):
        for part in (changed, extra):
            if part is None or not fw.be.changed_size(part):
                continue
            state["joined"] += fw.be.changed_size(part)
            hit = fw.be.join(part, bw.table(c, cache), tr.tmask[c])
            if hit is not None and (state["best"] is None or hit[0] < state["best"][0]):
                state["best"] = (hit[0], c) + tuple(hit[1:])
        return state["best"] is not None and state["best"][0] <= max(lower, 0)

    fw.run(on_cut=at_cut)
    cache.clear()
    rec.update(forward_counts=fw.counts, backward_counts=bw.counts, changed_counts=fw.changed_counts,
               peak_states=max(fw.peak, bw.peak), peak_table_bytes=fw.peak_bytes + bw.peak_bytes,
               backward_seconds=t1 - t0, forward_seconds=time.time() - t1,
               split=sched[0] if len(set(sched)) == 1 else _runs(sched), split_values=sched,
               joined=state["joined"], steps=fw.steps_done + bw.steps_done, dense_cuts=bw.dense_cuts)
    best = state["best"]
    if best is not None and best[0] <= K:
        total, c, fkey, fval, bkey = best
        rec["value"], rec["join_cut"] = total, c
        if witness:
            t2 = time.time()
            rec["witness"] = sorted(fw.traceback(c, fkey, fval) + bw.traceback(c, bkey, total - fval))
            rec["traceback_seconds"] = time.time() - t2
    rec["seconds"] = time.time() - t0
    return rec


# ---------------------------------------------------------------------------------------------- distance
def default_order(p, seeds=2):
    """The column order the dense planner picks (any order gives the same exact value)."""
    from . import auto
    return auto.best_order(p, seeds=seeds)[0]


def distance(p, order=None, start=1, kmax=64, mode="mitm", use_h=True, witness=True, log=None, every=None,
             backend="hip", dense=None, split=None):
    """The exact distance by raising the threshold from `start` (a lower bound): a threshold where no observable
    has a solution proves the next one.  Returns {"distance", "witness", "runs" (summaries per threshold and
    observable), "lower_bound", "order_cut_rank", "seconds"}; distance None if none is <= kmax."""
    t0 = time.time()
    order = default_order(p) if order is None else list(order)
    trellises = [Trellis.for_query(p, i, order) for i in range(p.nobs)]
    live = [(i, t) for i, t in enumerate(trellises) if t.feasible]
    out = {"distance": None, "witness": None, "runs": [], "lower_bound": start,
           "order_cut_rank": max((max(t.width) for _, t in live), default=0)}
    for K in range(start, kmax + 1):
        found = None
        for i, tr in live:
            r = query(tr, K, mode=mode, use_h=use_h, witness=witness, lower=K, every=every, backend=backend,
                      dense=dense, split=split(K) if callable(split) else split)
            r["obs"] = i
            out["runs"].append(summary(r))
            if log:
                log("  K=%d obs %d: value %s peak %d states %.2fs" % (K, i, r["value"], r["peak_states"],
                                                                       r["seconds"]))
            if r["value"] is not None:
                found = r                       # every lower threshold failed, so this value (<= K) is the least
                break
        if found is not None:
            out["distance"] = found["value"]
            if witness:
                cols = found["witness"]
                syn = lg = 0
                for j in cols:
                    syn ^= p.h[j]
                    lg ^= p.l[j]
                if syn or not lg or len(cols) != found["value"] or len(set(cols)) != len(cols):
                    raise AssertionError("witness check failed")
                out["witness"] = cols
            break
        out["lower_bound"] = K + 1
    out["seconds"] = time.time() - t0
    return out


def summary(r):
    """A query record with its per-cut lists replaced by their totals."""
    out = {}
    for k, v in r.items():
        if k.endswith("_counts"):
            out[k[:-len("_counts")] + "_states_total"] = int(sum(v))
        elif k not in ("witness", "split_values"):
            out[k] = v
    return out
