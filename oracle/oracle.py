"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).

CPU restatements of the reference's exact distance evaluators, used to check the HIP product path:

  run_plan_numpy(plan)          numpy executor of a Plan (cdist/kernel.py:72-91), extended to column weights
  run_plan_c(plan, impl)        the op-stream loop of cdist/kernel.py:94-138 over the C kernels of
                                impl="oracle" (liboracle.so, this project's scalar restatement) or
                                impl="ref" (oracle/_ref/libviterbi.so, the reference's viterbi.c compiled as is)
  brute(h, l, obs_mask, w)      exhaustive minimum over column subsets (tests/test_dp.py:12-23)
  brute_affine(columns, w)      all target minima (tests/test_tree_dp.py:44-54)
  tree joins, sparse steps, MITM tables: ctypes bindings of the same libraries

Pinning: tests/test_oracle_*.py check every function here against the reference's compiled C (_ref), brute force,
the chain closed form and the golden fixtures of tests/golden.  Nothing in tropidist_amd/ imports this module.
"""
from __future__ import annotations

import ctypes
import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 255
_LIBS = {}


def _load(path):
    if path not in _LIBS:
        if not os.path.exists(path):
            raise FileNotFoundError("%s is not built (make -C oracle)" % path)
        _LIBS[path] = ctypes.CDLL(path)
    return _LIBS[path]


def have_ref():
    return os.path.exists(os.path.join(HERE, "_ref", "libviterbi.so"))


def olib():
    L = _load(os.path.join(HERE, "liboracle.so"))
    if not getattr(L, "_typed", False):
        vp, u64, i64, ci = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int
        for name in ("or_dp_update",):
            getattr(L, name).argtypes = [vp, u64, u64]
        L.or_dp_update_w.argtypes = [vp, u64, u64, ci]
        L.or_dp_expand_update.argtypes = [vp, u64, u64]
        L.or_dp_expand_update_w.argtypes = [vp, u64, u64, ci]
        L.or_dp_select.argtypes = [vp, u64, ci, ci]
        L.or_dp_select.restype = u64
        L.or_dp_fill.argtypes = [vp, u64, ctypes.c_uint8]
        L.or_dp_hist.argtypes = [vp, u64, vp]
        L.or_run_ops.argtypes = [vp, i64, vp, ci]
        L.or_run_ops.restype = ci
        for b in (8, 16, 32, 64):
            f = getattr(L, "or_tree_join_u%d" % b)
            f.argtypes = [vp] * 5 + [ctypes.c_uint] * 2 + [u64] * 3
        L.or_sp_candidates.argtypes = [vp, i64, u64, u64, u64, ci, ci, vp, vp, i64, u64, ci, ci, ci, vp, vp, ci]
        L.or_sp_candidates.restype = i64
        L.or_sp_merge.argtypes = [vp, i64, vp, i64, u64, u64, ci, ci, ci, vp, vp, i64, u64, ci, ci, vp, vp,
                                  ctypes.POINTER(i64)]
        L.or_sp_merge.restype = i64
        L.or_sp_join.argtypes = [vp, i64, vp, i64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.or_sp_join.restype = ci
        L.or_sp_hist.argtypes = [vp, i64, vp]
        L.or_tab_create.argtypes = [ci, ci]
        L.or_tab_create.restype = vp
        L.or_tab_free.argtypes = [vp]
        L.or_tab_count.argtypes = [vp]
        L.or_tab_count.restype = i64
        L.or_tab_build.argtypes = [vp, ci, vp, vp, vp, ci]
        L.or_tab_lookup.argtypes = [vp, u64, u64, u64, ci]
        L.or_tab_lookup.restype = ci
        L._typed = True
    return L


def rlib(name):
    """The reference's own C kernel compiled into oracle/_ref (name: viterbi, tree_dp, sparse_dp, gmitm), with
    the argument types of its ctypes bindings (cdist/kernel.py:52-69, tree_dp.py:221-232, sparse_dp.py:484-499,
    search.py:54-69)."""
    L = _load(os.path.join(HERE, "_ref", "lib%s.so" % name))
    if getattr(L, "_typed", False):
        return L
    vp, u64, i64, ci = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int
    if name == "viterbi":
        L.dp_update.argtypes = [vp, u64, u64]
        L.dp_update_w.argtypes = [vp, u64, u64, ci]
        L.dp_expand_update.argtypes = [vp, u64, u64]
        L.dp_select.argtypes = [vp, u64, ci, ci]
        L.dp_select.restype = u64
        L.dp_fill.argtypes = [vp, u64, ctypes.c_uint8]
        L.dp_hist.argtypes = [vp, u64, vp]
        L.dp_num_threads.restype = ci
    elif name == "tree_dp":
        for b in (8, 16, 32, 64):
            f = getattr(L, "tree_join_u%d" % b)
            f.argtypes = [vp] * 5 + [ctypes.c_uint] * 2 + [u64] * 3
        L.tree_move.argtypes = [vp, vp, ctypes.c_size_t]
    elif name == "sparse_dp":
        L.sp_candidates.argtypes = [vp, i64, u64, u64, u64, ci, ci, vp, vp, i64, u64, ci, ci, ci, vp, vp, ci]
        L.sp_candidates.restype = i64
        L.sp_merge.argtypes = [vp, i64, vp, i64, u64, u64, ci, ci, ci, vp, vp, i64, u64, ci, ci, vp, vp,
                               ctypes.POINTER(i64)]
        L.sp_merge.restype = i64
        L.sp_join.argtypes = [vp, i64, vp, i64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.sp_join.restype = ci
        L.sp_hist.argtypes = [vp, i64, vp]
    elif name == "gmitm":
        L.tab_create.restype = vp
        L.tab_create.argtypes = [ci, ci]
        L.tab_free.argtypes = [vp]
        L.tab_count.restype = i64
        L.tab_count.argtypes = [vp]
        L.tab_build.argtypes = [vp, ci, vp, vp, vp, ci]
        L.gm_set_cover.argtypes = [ci, ci]
        L.gm_set_kernel.argtypes = [ci, vp, vp, vp, vp]
        L.gm_set_kernel_mode.argtypes = [ci]
        L.gm_set_full_labels.argtypes = [vp, ci]
        L.gm_set_pattern.argtypes = [vp, vp, ci, ci]
        L.gm_set_ledger.argtypes = [ctypes.c_char_p, ci, ci]
        L.gm_search.restype = ci
        L.gm_search.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, vp, vp, vp, ci, vp, ci, vp, vp, vp, ci,
                                ci, ci, ctypes.c_double, vp, vp, ctypes.POINTER(ci), ci, ci, vp, ctypes.POINTER(ci)]
    L._typed = True
    return L


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def final_value(entry):
    """cdist/kernel.py:18-25: 255 of a feasible plan means a cost >= 255."""
    v = int(entry)
    if v == INF:
        raise OverflowError("the distance is at least %d, beyond the 8-bit entries of the dense DP" % INF)
    return v


def _weight(op):
    return op[2] if len(op) > 2 else 1


def run_plan_numpy(plan):
    """cdist/kernel.py:72-91 with weighted updates: None if infeasible, OverflowError at 255."""
    if not plan.feasible:
        return None
    T = np.zeros(1, dtype=np.uint8)
    for op in plan.ops:
        if op[0] == "expand":
            T = np.concatenate([T, np.full(T.size, INF, dtype=np.uint8)])
        elif op[0] == "update":
            other = T[np.arange(T.size, dtype=np.int64) ^ op[1]].astype(np.int64) + _weight(op)
            T = np.minimum(T, np.minimum(other, INF)).astype(np.uint8)
        else:
            _, pos, target = op
            T = T[((np.arange(T.size, dtype=np.int64) >> pos) & 1) == target]
    assert T.size == 1
    return final_value(T[0])


def run_plan_c(plan, impl="oracle"):
    """The run_plan loop (cdist/kernel.py:94-138) on CPU C kernels: impl "oracle" runs liboracle's or_run_ops over
    the encoded op stream; impl "ref" drives the reference's own libviterbi.so op by op (unit weights call
    dp_update / dp_expand_update, weighted updates dp_update_w; a weighted expand-update has no reference kernel and
    is done as dp_fill + dp_update_w on the doubled table, which computes the same new half)."""
    if not plan.feasible:
        return None
    buf = np.empty(1 << plan.peak_bits, dtype=np.uint8)
    if impl == "oracle":
        v = olib().or_run_ops(ptr(plan.encode()), len(plan.ops), ptr(buf), plan.peak_bits)
        if v < 0:
            raise ValueError("malformed op stream")
        return final_value(v)
    L = rlib("viterbi")
    base = buf.ctypes.data
    buf[0] = 0
    n, i, ops = 1, 0, plan.ops
    while i < len(ops):
        op = ops[i]
        if op[0] == "expand":
            nxt = ops[i + 1] if i + 1 < len(ops) else None
            if nxt is not None and nxt[0] == "update" and (nxt[1] >> (n.bit_length() - 1)) == 1:
                if _weight(nxt) == 1:
                    L.dp_expand_update(base, n, nxt[1] ^ n)
                else:
                    L.dp_fill(base + n, n, INF)
                    L.dp_update_w(base, 2 * n, nxt[1], _weight(nxt))
                n <<= 1
                i += 2
                continue
            L.dp_fill(base + n, n, INF)
            n <<= 1
        elif op[0] == "update":
            if _weight(op) == 1:
                L.dp_update(base, n, op[1])
            else:
                L.dp_update_w(base, n, op[1], _weight(op))
        else:
            n = L.dp_select(base, n, op[1], op[2])
        i += 1
    assert n == 1
    return final_value(buf[0])


def brute(h, l, obs_mask, w=None):
    """Least total cost of a column set with H e = 0 and odd parity of (L e) & obs_mask; None if none
    (tests/test_dp.py:12-23, extended to costs: all subsets are scanned)."""
    n = len(h)
    best = None
    for size in range(1, n + 1):
        for sup in itertools.combinations(range(n), size):
            s = lg = 0
            for j in sup:
                s ^= h[j]
                lg ^= l[j]
            if s == 0 and bin(lg & obs_mask).count("1") & 1:
                c = size if w is None else sum(w[j] for j in sup)
                if best is None or c < best:
                    best = c
        if w is None and best is not None:
            return best
    return best


def brute_affine(columns, w=None):
    """target -> least cost of a column set summing to it (tests/test_tree_dp.py:44-54)."""
    result = {}
    for bits in itertools.product((0, 1), repeat=len(columns)):
        t = 0
        c = 0
        for j, (b, col) in enumerate(zip(bits, columns)):
            if b:
                t ^= col
                c += 1 if w is None else w[j]
        if t not in result or c < result[t]:
            result[t] = c
    return result


# ---------------------------------------------------------------------------------------------- tree DP
def tree_join(item, left, right, out, lm, rm, parent_bits, generators, start, stop, infinity, impl="oracle"):
    """One Gray-counter range of a join on numpy arrays (cdist/tree_dp.c:13-39): impl "oracle" (liboracle) or
    "ref" (the reference's libtree_dp.so)."""
    if impl == "oracle":
        f = getattr(olib(), "or_tree_join_u%d" % (8 * item))
    else:
        f = getattr(rlib("tree_dp"), "tree_join_u%d" % (8 * item))
    f(left.ctypes.data, right.ctypes.data, out.ctypes.data, lm.ctypes.data, rm.ctypes.data, parent_bits,
      generators, start, stop, infinity)


def tree_run(program, target=0, impl="oracle", chunk=1 << 20):
    """CompiledTreeDP.run restated on the CPU (cdist/tree_dp.py:401-505): the same postorder arena schedule with
    numpy tables and the C joins of `impl`.  Uses only the compiled host data of the product's CompiledTreeDP."""
    e0 = program.solver.solve(target)
    if e0 is None:
        return None
    if not program.nodes:
        return 0
    item, inf = program.itemsize, program.infinity
    arena = np.empty(program.peak_live_entries, dtype=program.dtype)
    maxj = max((j.join_dimension for j in program.joins.values()), default=0)
    maps = np.zeros((2, max(1, maxj)), dtype=np.uint64)
    stack, live = [], 0
    for index in program.order:
        node = program.nodes[index]
        size = 1 << len(program.interfaces[index])
        out = arena[live:live + size]
        out.fill(inf)
        if node.leaf is not None:
            bit = (e0 >> node.leaf) & 1
            out[0] = bit if program.columns[node.leaf] else 0
            if size == 2:
                out[1] = bit ^ 1
            stack.append((index, live, size))
            live += size
            continue
        second, first = stack.pop(), stack.pop()
        left, right = (first, second) if first[0] == node.left else (second, first)
        join = program.joins[index]
        gens = join.parent_maps + join.free_maps
        for b, g in enumerate(gens):
            maps[0, b] = g & ((1 << join.left_rank) - 1)
            maps[1, b] = g >> join.left_rank
        lt = arena[left[1]:left[1] + left[2]]
        rt = arena[right[1]:right[1] + right[2]]
        for begin in range(0, join.work, chunk):
            tree_join(item, lt, rt, out, maps[0], maps[1], join.parent_rank, join.join_dimension, begin,
                      min(begin + chunk, join.work), inf, impl)
        dest = first[1]
        arena[dest:dest + size] = out.copy()
        live = dest + size
        stack.append((index, dest, size))
    v = int(arena[0])
    return None if v == inf else v


# ---------------------------------------------------------------------------------------------- sparse tables
KEYMASK56 = (1 << 56) - 1


class OracleTables:
    """CPU sparse tables in the reference's C format (one sorted uint64 per state, key | value << 56, up to 56
    slots; cdist/sparse_dp.py:506-605) over liboracle's or_sp_candidates / or_sp_merge / or_sp_join (impl
    "oracle") or the reference's own libsparse_dp.so (impl "ref").  Implements the backend interface the product's
    Sweep and query use, so the same host logic can run on the CPU as a checker."""
    name = "oracle"
    impl = "oracle"

    def __init__(self, tr):
        if tr.nslots > 56:
            raise ValueError("the C tables hold at most 56 row slots (this trellis needs %d)" % tr.nslots)
        self.tr = tr
        self.nbits = max(1, tr.nslots)
        if self.impl == "oracle":
            L = olib()
            self.cand, self.merge, self.joinf = L.or_sp_candidates, L.or_sp_merge, L.or_sp_join
        else:
            L = rlib("sparse_dp")
            self.cand, self.merge, self.joinf = L.sp_candidates, L.sp_merge, L.sp_join
        self._bounds = {}

    @staticmethod
    def start():
        return np.zeros(1, dtype=np.uint64)

    @staticmethod
    def size(tab):
        return tab.count() if isinstance(tab, OracleDense) else tab.size

    @staticmethod
    def nbytes(tab):
        return tab.T.nbytes if isinstance(tab, OracleDense) else tab.nbytes

    @staticmethod
    def arrays(tab):
        if isinstance(tab, OracleDense):
            tab = tab.to_sparse()
        return tab & np.uint64(KEYMASK56), (tab >> np.uint64(56)).astype(np.uint8)

    @staticmethod
    def from_arrays(keys, vals):
        x = np.asarray(keys, dtype=np.uint64) | (np.asarray(vals).astype(np.uint64) << np.uint64(56))
        return x[np.argsort(x & np.uint64(KEYMASK56))]

    @staticmethod
    def lookup(tab, s):
        if isinstance(tab, OracleDense):
            return tab.lookup(s)
        keys = tab & np.uint64(KEYMASK56)
        i = int(np.searchsorted(keys, np.uint64(s)))
        if i < keys.size and int(keys[i]) == s:
            return int(tab[i] >> np.uint64(56))
        return INF

    def lookup_many(self, tab, keys):
        return np.array([self.lookup(tab, int(k)) for k in keys], dtype=np.uint8)

    def bound(self, forward):
        if forward not in self._bounds:
            self._bounds[forward] = [(np.array(m, dtype=np.uint64), np.array(w, dtype=np.int64), len(m), e)
                                     for m, w, e in self.tr.bound_data(forward)]
        return self._bounds[forward]

    def step(self, tab, upd, sel_mask, sel_bits, cap, K, tmask, hd, keep_below=None, want_changed=False):
        n = tab.size
        if hd is None:
            ncls, mp, wp, extra, use_h = 0, None, None, 0, 0
        else:
            m_arr, w_arr, ncls, extra = hd
            mp, wp, use_h = ptr(m_arr), ptr(w_arr), 1
        cand = np.empty(max(1, n), dtype=np.uint64)
        tmp = np.empty(max(1, n), dtype=np.uint64)
        m = self.cand(ptr(tab), n, upd, sel_mask, sel_bits, cap, ncls, mp, wp, extra, tmask, K, use_h, self.nbits,
                      ptr(cand), ptr(tmp), 1)
        out = np.empty(max(1, n + m), dtype=np.uint64)
        ch = np.empty(max(1, m), dtype=np.uint64) if want_changed else None
        nch = ctypes.c_int64(0)
        no = self.merge(ptr(tab), n, ptr(cand), m, sel_mask, sel_bits, cap, keep_below or 0, ncls, mp, wp, extra,
                        tmask, K, use_h, ptr(out), ptr(ch), ctypes.byref(nch))
        return out[:no].copy(), (ch[:nch.value].copy() if want_changed else None)

    @staticmethod
    def changed_size(ch):
        return ch.size

    @staticmethod
    def persistent(tab):
        return OracleDense(tab.slots, tab.T.copy()) if isinstance(tab, OracleDense) else tab.copy()

    @staticmethod
    def truncate(tab, below):
        return tab[(tab >> np.uint64(56)) < np.uint64(below)]

    def join(self, ch, btab, tmask):
        if not ch.size:
            return None
        if isinstance(btab, OracleDense):
            return btab.join(ch, tmask)
        fb, bb = ctypes.c_uint64(0), ctypes.c_uint64(0)
        tot = self.joinf(ptr(ch), ch.size, ptr(btab), btab.size, tmask, ctypes.byref(fb), ctypes.byref(bb))
        if tot < 0:
            return None
        return int(tot), fb.value & KEYMASK56, fb.value >> 56, bb.value & KEYMASK56

    @staticmethod
    def hist(tab):
        return np.bincount((tab >> np.uint64(56)).astype(np.int64), minlength=256)


    # ---- the dense form (cdist/sparse_dp.py:221-289, 680-707)
    @staticmethod
    def is_dense(tab):
        return isinstance(tab, OracleDense)

    @staticmethod
    def to_dense(tab, slots):
        D = OracleDense(slots, np.full(1 << len(slots), INF, dtype=np.uint8))
        keys, vals = tab & np.uint64(KEYMASK56), (tab >> np.uint64(56)).astype(np.uint8)
        D.T[D.compact(keys)] = vals
        return D

    @staticmethod
    def from_dense(D):
        return OracleTables.from_arrays(*D.arrays())

    def dense_step(self, D, starts, ends, upd, cap, K, tmask, hd):
        D = OracleDense(D.slots, D.T.copy())
        for s, _ in sorted(starts):
            D.insert_slot(s)
        cm = D.cmask(upd)
        if cm:
            olib().or_dp_update(ptr(D.T), D.T.size, cm)
        for s, t in ends:
            D.select_slot(s, t)
        D.prune(cap, K, tmask, hd)
        return D


class OracleDense:
    """A dense table of the sweeps' fallback, restated from cdist/sparse_dp.py:221-289: T[i] for the compact index
    i of a state (bit k of i = the state's bit at slots[k], slots ascending), INF = absent; the update is
    liboracle's restatement of viterbi.c's dp_update."""

    def __init__(self, slots, T):
        self.slots, self.T = list(slots), T

    def compact(self, keys):
        keys = np.asarray(keys, dtype=np.uint64)
        out = np.zeros(keys.shape, dtype=np.int64)
        for k, s in enumerate(self.slots):
            out |= ((keys >> np.uint64(s)) & np.uint64(1)).astype(np.int64) << k
        return out

    def expand(self, idx):
        out = np.zeros(idx.shape, dtype=np.uint64)
        for k, s in enumerate(self.slots):
            out |= ((idx >> k) & 1).astype(np.uint64) << np.uint64(s)
        return out

    def cmask(self, mask):
        return sum(1 << k for k, s in enumerate(self.slots) if (int(mask) >> s) & 1)

    def count(self):
        return int(np.count_nonzero(self.T != INF))

    def arrays(self):
        idx = np.flatnonzero(self.T != INF)
        return self.expand(idx), self.T[idx]

    def to_sparse(self):
        return OracleTables.from_arrays(*self.arrays())

    def lookup(self, s):
        if int(s) & ~sum(1 << x for x in self.slots):
            return INF
        return int(self.T[int(self.compact(np.array([s], dtype=np.uint64))[0])])

    def insert_slot(self, slot):                 # cdist/sparse_dp.py:262-267
        pos = sum(1 for x in self.slots if x < slot)
        A = self.T.reshape(-1, 1 << pos)
        self.T = np.concatenate([A, np.full_like(A, INF)], axis=1).reshape(-1)
        self.slots.insert(pos, slot)

    def select_slot(self, slot, target):         # cdist/sparse_dp.py:269-273
        pos = self.slots.index(slot)
        self.T = np.ascontiguousarray(self.T.reshape(-1, 2, 1 << pos)[:, target, :]).reshape(-1)
        self.slots.pop(pos)

    def prune(self, cap, K, tmask, hd):          # cdist/sparse_dp.py:275-289
        T = self.T
        T[(T > cap) & (T != INF)] = INF
        if hd is None:
            return
        masks, weights, extra = hd
        u = np.arange(T.size, dtype=np.uint64) ^ np.uint64(self.cmask(tmask))
        tot = np.full(T.size, int(extra), dtype=np.int64)
        for m, w in zip(masks, weights):
            tot += np.int64(int(w)) * np.bitwise_count(u & np.uint64(self.cmask(m))).astype(np.int64)
        h = (tot + (720720 - 1)) // 720720
        T[(T.astype(np.int64) + h > K) & (T != INF)] = INF

    def join(self, ch, tmask):                   # cdist/sparse_dp.py:420-444 over a dense table
        ck, cv = ch & np.uint64(KEYMASK56), (ch >> np.uint64(56)).astype(np.int64)
        q = ck ^ np.uint64(tmask)
        bv = np.array([self.lookup(int(x)) for x in q], dtype=np.int64)
        hit = bv != INF
        if not hit.any():
            return None
        tot = cv[hit] + bv[hit]
        i = int(np.argmin(tot))                  # ch is sorted by key: the least key among the least totals
        fi = int(np.flatnonzero(hit)[i])
        return int(tot[i]), int(ck[fi]), int(cv[fi]), int(q[fi])


class RefTables(OracleTables):
    """OracleTables over the reference's compiled sparse_dp.c."""
    name = "ref"
    impl = "ref"


# ---------------------------------------------------------------------------------------------- MITM tables
def tab_build(zh, zl, lg, k, log2slots, wide, impl="oracle"):
    """Build a table with liboracle (impl "oracle") or the reference's libgmitm.so (impl "ref"); returns
    (handle, count, free function)."""
    zh, zl, lg = (np.ascontiguousarray(a, dtype=np.uint64) for a in (zh, zl, lg))
    if impl == "oracle":
        L = olib()
        h = L.or_tab_create(log2slots, wide)
        L.or_tab_build(h, zh.size, ptr(zh), ptr(zl), ptr(lg), k)
        return h, L.or_tab_count(h), L.or_tab_free
    L = rlib("gmitm")
    h = L.tab_create(log2slots, wide)
    L.tab_build(h, zh.size, ptr(zh), ptr(zl), ptr(lg), k)
    return h, L.tab_count(h), L.tab_free


def tab_lookup(handle, zh, zl, lg, R):
    return olib().or_tab_lookup(handle, int(zh), int(zl), int(lg), int(R))


def brute_tab_lookup(zh, zl, lg, k, qzh, qzl, qlg, R, wide):
    """Least size <= min(k, R) of a column subset whose syndrome key matches the query's (compact: bits 16-62 of
    zl; wide: all of zl) with a different label (compact: low 16 bits), among subsets that share the query's
    probe chain start is not required -- the exhaustive set of matching entries; -1 if none."""
    fp = 0x7FFFFFFFFFFF0000
    best = -1
    n = len(zh)
    for size in range(0, min(k, R) + 1):
        for sup in itertools.combinations(range(n), size):
            a = b = g = 0
            for j in sup:
                a ^= int(zh[j])
                b ^= int(zl[j])
                g ^= int(lg[j])
            if wide:
                ok = b == int(qzl) and g != int(qlg)
            else:
                ok = (b & fp) == (int(qzl) & fp) and (g & 0xFFFF) != (int(qlg) & 0xFFFF)
            if ok and a == int(qzh):
                return size
    return best


# ---------------------------------------------------------------------------------------------- sparse trellis
class SparseTrellis:
    """The sparse trellis of one query, restated from cdist/sparse_dp.py:63-145 for checking the native builder:
    rows = the check rows over positions in minimal span form plus the reduced target row (last, target 1); every
    row gets a slot by colouring the closed intervals [start, end] in order of start, reusing the lowest slot whose
    row ended before the start; per position p: col[p] = slots of the rows containing p, fsel_mask[p] = slots of rows
    ending at p (fsel_bits: the target row's), bsel_mask[p] / bsel_bits[p] the same for rows starting at p; per cut
    c: active rows are those with start < c <= end, tmask[c] = the target row's slot while it is active.  Raises
    ValueError beyond max_slots.  The span-form routines are the dense planner's (pinned by the recorded plans)."""

    def __init__(self, h, lrow, order, nrows, max_slots=120):
        import heapq
        from tropidist_amd import trellis as dense
        self.order, self.N = list(order), len(order)
        hrows = dense.minimal_span_form(dense.rows_from_columns(h, self.order, nrows))
        target = dense.reduce_target_row(lrow, hrows)
        self.feasible = target != 0
        if not self.feasible:
            return
        rows = list(hrows) + [target]
        lo = [(r & -r).bit_length() - 1 for r in rows]
        hi = [r.bit_length() - 1 for r in rows]
        slot, pool, ending, used = [None] * len(rows), [], [], 0
        for r in sorted(range(len(rows)), key=lambda x: (lo[x], x)):
            while ending and ending[0][0] < lo[r]:
                heapq.heappush(pool, heapq.heappop(ending)[1])
            if pool:
                slot[r] = heapq.heappop(pool)
            else:
                slot[r], used = used, used + 1
            heapq.heappush(ending, (hi[r], slot[r]))
        if used > max_slots:
            raise ValueError("trellis needs %d bits per state (at most %d supported)" % (used, max_slots))
        self.nslots, self.rows, self.slot, self.lo, self.hi = used, rows, slot, lo, hi
        N, last = self.N, len(rows) - 1
        self.col = [0] * N
        self.fsel_mask, self.fsel_bits = [0] * N, [0] * N
        self.bsel_mask, self.bsel_bits = [0] * N, [0] * N
        for r, row in enumerate(rows):
            bit = 1 << slot[r]
            for p in range(lo[r], hi[r] + 1):
                if (row >> p) & 1:
                    self.col[p] |= bit
            self.fsel_mask[hi[r]] |= bit
            self.bsel_mask[lo[r]] |= bit
            if r == last:
                self.fsel_bits[hi[r]] |= bit
                self.bsel_bits[lo[r]] |= bit
        self.active_rows = [[r for r in range(len(rows)) if lo[r] < c <= hi[r]] for c in range(N + 1)]
        self.width = [len(a) for a in self.active_rows]
        self.tmask = [(1 << slot[last]) if lo[last] < c <= hi[last] else 0 for c in range(N + 1)]

    def bound_data(self, forward, Q=720720):
        """The lookahead bound per cut (cdist/sparse_dp.py:148-192): (slot masks by weight ascending, weights,
        extra), computed position by position with Python ints."""
        rows, last, out = self.rows, len(self.rows) - 1, []
        for c in range(self.N + 1):
            rel = list(self.active_rows[c])
            lone = (forward and self.lo[last] >= c) or (not forward and self.hi[last] < c)
            if lone:
                rel.append(last)
            if not rel:
                out.append(((), (), 0))
                continue
            window = range(c, max(self.hi[r] for r in rel) + 1) if forward else range(min(self.lo[r] for r in rel), c)
            cover = {p: sum((rows[r] >> p) & 1 for r in rel) for p in window}
            by_weight, extra = {}, 0
            for r in rel:
                most = max(cover[p] for p in window if (rows[r] >> p) & 1)
                w = Q // most
                if lone and r == last:
                    extra = w
                else:
                    by_weight[w] = by_weight.get(w, 0) | (1 << self.slot[r])
            ws = sorted(by_weight)
            out.append((tuple(by_weight[w] for w in ws), tuple(ws), extra))
        return out


def bound_value(state, tmask, hd, Q=720720):
    """The lookahead lower bound of one state at a cut: ceil((extra + sum_k w_k |(s ^ tmask) & mask_k|) / Q)."""
    masks, weights, extra = hd
    u = int(state) ^ int(tmask)
    tot = int(extra) + sum(int(w) * bin(u & int(m)).count("1") for m, w in zip(masks, weights))
    return -(-tot // Q)


# ---------------------------------------------------------------------------------------------- weight-bounded search
def have_gmitm():
    return os.path.exists(os.path.join(HERE, "_ref", "libgmitm.so"))


def gm_search_ref(S, K, tables, main, rank, anchors, use_ex=True, enum_all=False, task_depth=1, sol_cap=100000,
                  cover_gate=0, ascent=1, kernel=None, kernel_mode=False, pattern=None, dmax=40):
    """The reference's own search (gmitm.c gm_search, compiled from source into oracle/_ref/libgmitm.so) on the
    arrays of a tropidist_amd.search.Searcher S.  tables: (zh, zl, lg, k, log2slots, wide) per table, built here as
    the reference's CPU tables with the same hashes.  Returns the record fields the device search reports: status
    code, best, sorted solutions, nodes per depth and the eight statistics."""
    L = rlib("gmitm")
    built = [tab_build(zh, zl, lg, k, log2, wide, impl="ref") for zh, zl, lg, k, log2, wide in tables]
    keep = [(np.ascontiguousarray(t[0], dtype=np.uint64), np.ascontiguousarray(t[1], dtype=np.uint64))
            for t in tables]
    nt = len(tables)
    handles = (ctypes.c_void_p * max(nt, 1))(*[b[0] for b in built])
    tzh = (ctypes.c_void_p * max(nt, 1))(*[k[0].ctypes.data for k in keep])
    tzl = (ctypes.c_void_p * max(nt, 1))(*[k[1].ctypes.data for k in keep])
    L.gm_set_cover(int(cover_gate), int(ascent))
    if kernel is not None:
        nks, a, b, c, d = kernel
        L.gm_set_kernel(nks, ptr(a), ptr(b), ptr(c), ptr(d))
    else:
        L.gm_set_kernel(0, None, None, None, None)
    L.gm_set_kernel_mode(1 if kernel_mode else 0)
    L.gm_set_full_labels(ptr(S.full_lg), S.lw if S.full_lg is not None else 0)
    pkeep = []
    if pattern is not None:
        cls = np.ascontiguousarray(pattern[0], dtype=np.int32)
        onpat = np.ascontiguousarray(pattern[1], dtype=np.uint8)
        pkeep = [cls, onpat]
        L.gm_set_pattern(ptr(cls), ptr(onpat), len(onpat), int(pattern[2]))
    else:
        L.gm_set_pattern(None, None, 0, 0)
    L.gm_set_ledger(None, 0, 1)
    rank = np.ascontiguousarray(rank, dtype=np.int32)
    anchors = np.ascontiguousarray(anchors, dtype=np.int32)
    nodes = np.zeros(dmax + 2, dtype=np.int64)
    stats = np.zeros(8, dtype=np.int64)
    stride = K + 2
    sol = np.zeros((max(sol_cap, 1), stride), dtype=np.int32)
    best, nsol = ctypes.c_int(-1), ctypes.c_int(0)
    try:
        rc = L.gm_search(S.W, S.p.ncols, S.p.nrows, S.wmax, ptr(S.col), ptr(S.lg), ptr(S.row_start),
                         ptr(S.row_cols), ptr(S.col_start), ptr(S.col_rows), nt, handles, tzh, tzl, main, ptr(rank),
                         len(anchors), ptr(anchors), ptr(S.ex_start) if use_ex else None,
                         ptr(S.ex_cols) if use_ex else None, K, 1 if enum_all else 0, task_depth, 0.0, ptr(nodes),
                         ptr(stats), ctypes.byref(best), sol_cap, stride, ptr(sol), ctypes.byref(nsol))
    finally:
        L.gm_set_kernel_mode(0)
        L.gm_set_kernel(0, None, None, None, None)
        L.gm_set_cover(0, 0)
        L.gm_set_full_labels(None, 0)
        L.gm_set_pattern(None, None, 0, 0)
        for h, _, free in built:
            free(h)
    del pkeep
    sols = sorted(tuple(sorted(int(x) for x in sol[i, 1:1 + int(sol[i, 0])])) for i in range(min(nsol.value, sol_cap)))
    return {"status": rc, "best": None if best.value < 0 else best.value, "solutions": sols,
            "nodes": [int(x) for x in nodes[:K + 2]], "stats": [int(x) for x in stats]}


def brute_all(h, l):
    """Minimum weight and all minimum-weight supports with zero syndrome and nonzero logical (exhaustive, small n;
    the check of tests/test_search.py:13-28)."""
    n = len(h)
    for w in range(1, n + 1):
        sols = []
        for sup in itertools.combinations(range(n), w):
            s = lg = 0
            for j in sup:
                s ^= h[j]
                lg ^= l[j]
            if s == 0 and lg != 0:
                sols.append(sup)
        if sols:
            return w, sols
    return None, []
