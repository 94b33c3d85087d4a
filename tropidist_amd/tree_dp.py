"""Exact dense min-plus DP on a binary tree of fault columns: compiled in C++, joined on the GPU.

The call surface is that of the reference's cdist/tree_dp.py: CompiledTreeDP(columns, tree).run(target) is the least
Hamming weight of e with M e = target (columns: nonnegative Python-int bitmasks; tree: an int leaf or a pair of trees
whose leaves are exactly range(len(columns)); None for no columns), with .preflight(**limits) reporting the exact
device arena and join work before anything is allocated; tree_min_distance_for_mask / preflight_tree_dp answer one
logical query; TreeDPResourceError(message, profile) is raised when a limit would be exceeded.

Compilation (libtdp's tdp_tree_compile) finds, at every node S, the interface I_S = im(M_S) /\\ im(M_rest) -- the
states a partial assignment of S's columns can leave for the rest -- and, at every join, the generator maps of the
compatible child-state pairs; the execution order is a postorder choosing, per join, the child order with the
smaller peak of live table entries.  Execution (tdp_tree_run) keeps every table of the postorder stack in one device
arena of exactly the preflight size: leaf tables, then per join a fill, the min-plus XOR join of the two child tables
on the GPU (tdp_tree_join, in chunks of Gray-counter ranges) and the move of the parent table down to its first
child's place.  Costs are unsigned integers of the smallest width holding ncols + 1, the infinity."""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

DEFAULT_MEMORY_BYTES = 256 << 20
DEFAULT_JOIN_WORK = 100_000_000
_CHUNK_PAIRS = 1 << 62          # Gray-counter pairs per join launch (a join is one launch unless this is lowered)


class TreeDPResourceError(RuntimeError):
    """A compilation, device memory, join work or time limit was exceeded; .profile holds the preflight record."""

    def __init__(self, message, profile=None):
        super().__init__(message)
        self.profile = profile


@dataclass(frozen=True)
class TreeNode:
    """A node of the flattened tree: a leaf (column index) or the node indices of its two children."""
    leaf: int | None
    left: int | None = None
    right: int | None = None


@dataclass(frozen=True)
class TreeJoin:
    """A join's shape and generator maps: generator g moves the left index by lmaps[g] and the right by rmaps[g];
    the first parent_rank generators also flip parent bit g, the others (the kernel of the child-state sum) leave
    the parent index alone."""
    left_rank: int
    right_rank: int
    parent_rank: int
    join_dimension: int
    lmaps: tuple = ()
    rmaps: tuple = ()

    @property
    def work(self):
        return 1 << self.join_dimension

    @property
    def kernel_rank(self):
        return self.join_dimension - self.parent_rank

    def _packed(self, lo, hi):
        return tuple(int(l) | (int(r) << self.left_rank) for l, r in zip(self.lmaps[lo:hi], self.rmaps[lo:hi]))

    @property
    def parent_maps(self):
        """Parent generators as one coefficient mask over (left basis, right basis)."""
        return self._packed(0, self.parent_rank)

    @property
    def free_maps(self):
        return self._packed(self.parent_rank, self.join_dimension)


def _is_index(x):
    return isinstance(x, int) and not isinstance(x, bool)


def _flatten(tree, ncols):
    """Postorder node list of a tree of column leaves.  The walk visits (node, right subtree, left subtree) in
    preorder with an explicit stack, and the reverse of that order is a postorder; at most 2 ncols - 1 visits are
    allowed, which stops cycles and repeated subtrees."""
    if ncols == 0:
        if tree is not None:
            raise ValueError("the empty problem requires tree=None")
        return ()
    visits, seen = [], bytearray(ncols)
    kids = []                                   # per visit: [left visit, right visit] (internal) or None (leaf)
    todo = [(tree, -1, 0)]
    while todo:
        item, parent, side = todo.pop()
        if len(visits) >= 2 * ncols - 1:
            raise ValueError("tree has too many nodes, repeated leaves, or a cycle")
        me = len(visits)
        visits.append(item)
        if parent >= 0:
            kids[parent][side] = me
        if _is_index(item):
            if not 0 <= item < ncols or seen[item]:
                raise ValueError("tree leaves must be each column index exactly once")
            seen[item] = 1
            kids.append(None)
        elif isinstance(item, (tuple, list)) and len(item) == 2:
            kids.append([-1, -1])
            todo.append((item[0], me, 0))       # popped after the right child
            todo.append((item[1], me, 1))
        else:
            raise ValueError("each tree node must be an integer leaf or a binary pair")
    if sum(seen) != ncols:
        raise ValueError("tree leaves must be each column index exactly once")
    last = len(visits) - 1
    nodes = []
    for v in range(last, -1, -1):                # postorder index of visit v is last - v
        if kids[v] is None:
            nodes.append(TreeNode(visits[v]))
        else:
            nodes.append(TreeNode(None, last - kids[v][0], last - kids[v][1]))
    return tuple(nodes)


def _cost_dtype(ncols):
    """The narrowest unsigned dtype holding ncols + 1 (the infinity; costs never exceed ncols)."""
    for dt in (np.uint8, np.uint16, np.uint32, np.uint64):
        if ncols + 1 <= np.iinfo(dt).max:
            return np.dtype(dt)
    raise ValueError("number of columns exceeds supported exact integer range")


def _words(values, nw):
    """Python ints as rows of nw little-endian uint64 words."""
    out = np.zeros((max(1, len(values)), nw), dtype=np.uint64)
    for i, v in enumerate(values):
        for k in range(nw):
            out[i, k] = (v >> (64 * k)) & 0xFFFFFFFFFFFFFFFF
    return out


def _int_of(words):
    return sum(int(w) << (64 * k) for k, w in enumerate(words))


def compile_storage_bound(columns, nodes):
    """Host bytes the C++ compiler may hold for this tree, from sizes alone: per node up to min(rank bound, leaves)
    vectors of span, outer span and interface, the column solver with its tags, and the join eliminations."""
    ncols = len(columns)
    bits = max((v.bit_length() for v in columns), default=0)
    rw = max(1, (bits + 63) // 64)
    rank_bound = min(ncols, bits)
    leaves, total = [], 0
    for nd in nodes:
        leaves.append(1 if nd.leaf is not None else leaves[nd.left] + leaves[nd.right])
        total += min(rank_bound, leaves[-1])
    cw = (ncols + 63) // 64 + 1
    return (8 * rw * ncols + 3 * 8 * rw * total + 8 * (rw + cw) * rank_bound
            + 16 * (rank_bound + 1) ** 2 + 160 * len(nodes) + 16384)


class _Solver:
    """Particular solutions e0 of M e0 = t from the compiled program (None when t is outside the column span)."""

    def __init__(self, program):
        self._p = program

    def solve(self, target):
        return self._p.particular_solution(target)


class CompiledTreeDP:
    """A binary tree realisation of M, reusable for any target.  max_compile_bytes bounds compile_storage_bound()
    before compiling; max_compile_seconds is checked between nodes of the compiler.  Neither compilation nor
    preflight allocates an exponential-size array."""

    def __init__(self, columns, tree, *, max_compile_bytes=DEFAULT_MEMORY_BYTES, max_compile_seconds=60.0):
        t0 = time.monotonic()
        self.columns = tuple(columns)
        if any(not _is_index(v) or v < 0 for v in self.columns):
            raise ValueError("columns must be nonnegative Python integers")
        self.ncols = len(self.columns)
        self.nodes = _flatten(tree, self.ncols)
        self.root = len(self.nodes) - 1
        dt = _cost_dtype(self.ncols)
        self.dtype, self.itemsize, self.infinity = dt.name, dt.itemsize, self.ncols + 1
        self.compile_storage_bound_bytes = compile_storage_bound(self.columns, self.nodes)
        if max_compile_bytes is not None and self.compile_storage_bound_bytes > max_compile_bytes:
            raise TreeDPResourceError(
                "conservative compilation bound %d bytes exceeds max_compile_bytes=%d"
                % (self.compile_storage_bound_bytes, max_compile_bytes),
                {"compile_storage_bound_bytes": self.compile_storage_bound_bytes})
        bits = max((v.bit_length() for v in self.columns), default=0)
        self._rw = max(1, (bits + 63) // 64)
        self._handle = None
        self._L = _lib.lib()
        n = len(self.nodes)
        leaf = np.array([-1 if nd.leaf is None else nd.leaf for nd in self.nodes] or [0], dtype=np.int32)
        left = np.array([-1 if nd.left is None else nd.left for nd in self.nodes] or [0], dtype=np.int32)
        right = np.array([-1 if nd.right is None else nd.right for nd in self.nodes] or [0], dtype=np.int32)
        cols = _words(self.columns, self._rw)
        status = ctypes.c_int(0)
        limit = -1.0 if max_compile_seconds is None else max(0.0, max_compile_seconds - (time.monotonic() - t0))
        h = self._L.tdp_tree_compile(cols.ctypes.data, self.ncols, self._rw, leaf.ctypes.data, left.ctypes.data,
                                     right.ctypes.data, n, limit, ctypes.byref(status))
        if not h:
            msg = self._L.tdp_last_error().decode(errors="replace")
            if status.value == _lib.TDP_EDEADLINE:
                raise TreeDPResourceError("tree compilation exceeded max_compile_seconds")
            raise AssertionError(msg) if "identity" in msg or "inside" in msg else ValueError(msg)
        self._handle = h
        info = (ctypes.c_int64 * 6)()
        _lib.check(self._L.tdp_tree_info(h, info), "tdp_tree_info")
        self.matrix_rank, self.peak_live_entries = int(info[0]), int(info[1])
        self._metadata_native = int(info[2])
        ranks = np.zeros(max(1, n), dtype=np.int32)
        order = np.zeros(max(1, n), dtype=np.int32)
        first = np.zeros(max(1, n), dtype=np.int32)
        shape = np.zeros((max(1, n), 4), dtype=np.int32)
        if n:
            _lib.check(self._L.tdp_tree_nodes(h, ranks.ctypes.data, order.ctypes.data, first.ctypes.data,
                                              shape.ctypes.data), "tdp_tree_nodes")
        self.interface_ranks = tuple(int(r) for r in ranks[:n])
        self.order = tuple(int(x) for x in order[:n])
        self.joins, self.execution_children = {}, {}
        for i, nd in enumerate(self.nodes):
            if nd.leaf is not None:
                continue
            lr, rr, pr, J = (int(x) for x in shape[i])
            lm = rm = ()
            if lr <= 63 and rr <= 63 and J <= 62:
                a = np.zeros(max(1, J), dtype=np.uint64)
                b = np.zeros(max(1, J), dtype=np.uint64)
                _lib.check(self._L.tdp_tree_maps(h, i, a.ctypes.data, b.ctypes.data), "tdp_tree_maps")
                lm, rm = tuple(int(x) for x in a[:J]), tuple(int(x) for x in b[:J])
            self.joins[i] = TreeJoin(lr, rr, pr, J, lm, rm)
            f = int(first[i])
            self.execution_children[i] = (f, nd.right if f == nd.left else nd.left)
        self.solver = _Solver(self)
        self._interfaces = None
        self.compilation_seconds = time.monotonic() - t0
        self.compiled_metadata_bytes = self._metadata_native + 200 * (len(self.nodes) + len(self.joins))

    def __del__(self):
        h, self._handle = getattr(self, "_handle", None), None
        if h:
            try:
                self._L.tdp_tree_free(h)
            except Exception:
                pass

    @property
    def interfaces(self):
        """Per node, a basis of its interface (Python ints), read from the compiled program on first use."""
        if self._interfaces is None:
            out = []
            for i, r in enumerate(self.interface_ranks):
                buf = np.zeros((max(1, r), self._rw), dtype=np.uint64)
                if r:
                    _lib.check(self._L.tdp_tree_interface(self._handle, i, buf.ctypes.data), "tdp_tree_interface")
                out.append(tuple(_int_of(row) for row in buf[:r]))
            self._interfaces = tuple(out)
        return self._interfaces

    def particular_solution(self, target):
        """A column set e0 (bitmask) with M e0 = target, or None if the target is outside the column span."""
        if target >> (64 * self._rw):
            return None
        t = _words([target], self._rw)[0]
        e0 = np.zeros((self.ncols + 63) // 64 + 1, dtype=np.uint64)
        ok = ctypes.c_int(0)
        if self._handle is None:
            return 0 if target == 0 else None
        _lib.check(self._L.tdp_tree_solve(self._handle, t.ctypes.data, e0.ctypes.data, ctypes.byref(ok)),
                   "tdp_tree_solve")
        return _int_of(e0) if ok.value else None

    def preflight(self, *, max_memory_bytes=DEFAULT_MEMORY_BYTES, max_join_work=DEFAULT_JOIN_WORK,
                  max_table_bits=None):
        """Exact device sizes, candidate-pair counts and the reasons a run with these limits would be refused
        (cdist/tree_dp.py:349-399 reports the same record).  `allowed` says nothing about a target's feasibility."""
        ranks = list(self.interface_ranks)
        max_rank = max(ranks, default=0)
        max_join = max((j.join_dimension for j in self.joins.values()), default=0)
        work = sum(j.work for j in self.joins.values())
        table_bytes = self.peak_live_entries * self.itemsize
        scratch_bytes = 16 * max_join            # the two uint64 generator maps of the widest join
        peak = table_bytes + scratch_bytes
        refuse = []
        if max_rank >= 63 or max_join >= 63:
            refuse.append("table or join index requires at least 63 bits")
        if table_bytes > np.iinfo(np.intp).max:
            refuse.append("arena exceeds the platform array index range")
        if max_table_bits is not None and max_rank > max_table_bits:
            refuse.append("interface rank %d exceeds max_table_bits=%d" % (max_rank, max_table_bits))
        if max_memory_bytes is not None and peak > max_memory_bytes:
            refuse.append("dense workspace %d bytes exceeds max_memory_bytes=%d" % (peak, max_memory_bytes))
        if max_join_work is not None and work > max_join_work:
            refuse.append("join work %d exceeds max_join_work=%d" % (work, max_join_work))
        profiles = []
        for i, j in self.joins.items():
            tables = (1 << j.left_rank) + (1 << j.right_rank) + (1 << j.parent_rank)
            profiles.append({"node": i, "left_rank": j.left_rank, "right_rank": j.right_rank,
                             "parent_rank": j.parent_rank, "kernel_rank": j.kernel_rank,
                             "join_dimension": j.join_dimension, "work": j.work,
                             "child_parent_table_bytes": self.itemsize * tables})
        return {
            "allowed": not refuse, "rejection_reasons": refuse,
            "ncols": self.ncols, "matrix_rank": self.matrix_rank,
            "dtype": self.dtype, "dtype_bytes": self.itemsize, "infinity": self.infinity,
            "edge_state_dimensions": ranks, "max_interface_rank": max_rank,
            "max_join_dimension": max_join, "join_work": work, "join_profiles": profiles,
            "largest_table_bytes": (1 << max_rank) * self.itemsize if self.nodes else 0,
            "peak_live_table_bytes": table_bytes, "arena_bytes": table_bytes,
            "scratch_bytes": scratch_bytes, "peak_dense_bytes": peak,
            "compiled_metadata_bytes": self.compiled_metadata_bytes,
            "compile_storage_bound_bytes": self.compile_storage_bound_bytes,
            "compilation_seconds": self.compilation_seconds,
            "memory_scope": "device arena and generator maps; the compiled host program reported separately",
            "threads": "gpu",
        }

    def run(self, target=0, *, max_memory_bytes=DEFAULT_MEMORY_BYTES, max_join_work=DEFAULT_JOIN_WORK,
            max_table_bits=None, max_seconds=60.0, backend="hip", stream=None):
        """(least weight of e with M e = target, or None when there is none; statistics).  backend "hip" (or "c",
        the reference's name for its native path) runs the joins on the GPU; the limits are checked before any
        device allocation and the time limit between launches."""
        if not _is_index(target) or target < 0:
            raise ValueError("target must be a nonnegative Python integer")
        if backend not in ("hip", "c"):
            raise ValueError("backend must be 'hip' (or its alias 'c')")
        t0 = time.monotonic()
        profile = self.preflight(max_memory_bytes=max_memory_bytes, max_join_work=max_join_work,
                                 max_table_bits=max_table_bits)
        e0 = self.particular_solution(target)
        base = {"seconds": 0.0, "join_work_done": 0, "allocated_dense_bytes": 0, "backend": "hip"}
        if e0 is None:
            return None, dict(profile, feasible=False, **base)
        if not profile["allowed"]:
            raise TreeDPResourceError("; ".join(profile["rejection_reasons"]), profile)
        if not self.nodes:
            return 0, dict(profile, feasible=True, **base)
        if max_seconds is not None and time.monotonic() - t0 > max_seconds:
            raise TreeDPResourceError("tree DP exceeded max_seconds; no exact result returned", profile)
        arena = _lib.device_empty(self.peak_live_entries * self.itemsize)
        ew = np.zeros((self.ncols + 63) // 64 + 1, dtype=np.uint64)
        for k in range(ew.size):
            ew[k] = (e0 >> (64 * k)) & 0xFFFFFFFFFFFFFFFF
        value = ctypes.c_uint64(0)
        stats = (ctypes.c_uint64 * 2)()
        left = -1.0 if max_seconds is None else max_seconds - (time.monotonic() - t0)
        rc = self._L.tdp_tree_run(self._handle, ew.ctypes.data, self.itemsize, arena.data_ptr(),
                                  self.peak_live_entries, min(_CHUNK_PAIRS, 1 << 62), left, ctypes.byref(value),
                                  stats, _lib.stream_handle(stream))
        if rc == _lib.TDP_EDEADLINE:
            raise TreeDPResourceError("tree DP exceeded max_seconds; no exact result returned", profile)
        _lib.check(rc, "tdp_tree_run")
        v = int(value.value)
        if v == self.infinity:
            raise AssertionError("a consistent affine system must have a finite root cost")
        return v, dict(profile, feasible=True, seconds=time.monotonic() - t0, join_work_done=int(stats[0]),
                       allocated_dense_bytes=profile["peak_dense_bytes"],
                       observed_peak_live_table_bytes=int(stats[1]) * self.itemsize, backend="hip")


def _validate_query(h_cols, l_cols, mask):
    if len(h_cols) != len(l_cols):
        raise ValueError("H and L must have the same number of columns")
    if not _is_index(mask) or mask < 0:
        raise ValueError("observable mask must be a nonnegative Python integer")
    for v in (*h_cols, *l_cols):
        if not _is_index(v) or v < 0:
            raise ValueError("columns must be nonnegative Python integers")


def _with_parity_row(h_cols, l_cols, mask):
    """The columns with one more row on top: the parity of the column's observables in `mask`; the query's target
    is that row alone."""
    top = max((v.bit_length() for v in h_cols), default=0)
    return top, tuple(h | ((bin(l & mask).count("1") & 1) << top) for h, l in zip(h_cols, l_cols))


def tree_min_distance_for_mask(h_cols, l_cols, mask, tree, *, max_compile_bytes=DEFAULT_MEMORY_BYTES,
                               max_compile_seconds=60.0, **limits):
    """min |e| subject to H e = 0 and odd parity of (L e) & mask, on the given tree (cdist/tree_dp.py:508-527)."""
    _validate_query(h_cols, l_cols, mask)
    top, cols = _with_parity_row(h_cols, l_cols, mask)
    program = CompiledTreeDP(cols, tree, max_compile_bytes=max_compile_bytes,
                             max_compile_seconds=max_compile_seconds)
    distance, info = program.run(1 << top, **limits)
    info["observable_mask"] = mask
    return distance, info


def preflight_tree_dp(h_cols, l_cols, mask, tree, *, max_compile_bytes=DEFAULT_MEMORY_BYTES,
                      max_compile_seconds=60.0, **limits):
    """The compiled query's preflight record and whether its target is reachable (cdist/tree_dp.py:530-546)."""
    _validate_query(h_cols, l_cols, mask)
    top, cols = _with_parity_row(h_cols, l_cols, mask)
    program = CompiledTreeDP(cols, tree, max_compile_bytes=max_compile_bytes,
                             max_compile_seconds=max_compile_seconds)
    return dict(program.preflight(**limits), feasible=program.particular_solution(1 << top) is not None,
                observable_mask=mask)
