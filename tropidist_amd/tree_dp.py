"""Exact dense min-plus DP on an arbitrary binary tree of fault columns, with the joins on the GPU.

Call surface of the reference's cdist/tree_dp.py: CompiledTreeDP(M, tree).run(t) minimises the Hamming weight
subject to M e = t; columns are nonnegative Python-int bitmasks; a tree is an int leaf or a pair of trees whose
leaves are exactly range(len(M)); None is the empty tree.  tree_min_distance_for_mask / preflight_tree_dp answer a
logical query; TreeDPResourceError is raised before any device allocation when a limit would be exceeded.

Compilation is host work (GF(2) bases, interfaces I_S = im(M_S) /\\ im(M_notS), the join generator maps and the
postorder arena schedule, cdist/tree_dp.py:245-347).  Execution allocates one device arena of the exact
preflight size, initialises leaf tables, runs each join as one GPU launch (tdp_tree_join) and compacts the
postorder stack on the device (tdp_tree_move).  The exact J identity
    J = rank(M_A) + rank(M_B) + rank(M_outside) - rank(M)
is checked for every join, as in the reference."""
from __future__ import annotations

import ctypes
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

DEFAULT_MEMORY_BYTES = 256 << 20
DEFAULT_JOIN_WORK = 100_000_000
_CHUNK_PAIRS = 1 << 62          # a join runs as one launch; smaller values split it into Gray-counter ranges
_popcount = getattr(int, "bit_count", lambda v: bin(v).count("1"))


class TreeDPResourceError(RuntimeError):
    """Compilation, dense memory, candidate-pair work, or time limit exceeded (cdist/tree_dp.py:50-55)."""

    def __init__(self, message, profile=None):
        super().__init__(message)
        self.profile = profile


def _basis(columns):
    """Independent vectors spanning the columns, reduced by their top bits."""
    piv = {}
    for v in columns:
        while v:
            t = v.bit_length() - 1
            if t not in piv:
                piv[t] = v
                break
            v ^= piv[t]
    return tuple(piv.values())


class _LinearMap:
    """Column elimination that remembers, for each pivot, which input columns sum to it; columns that reduce to
    zero give kernel vectors (cdist/tree_dp.py:70-99)."""

    def __init__(self, columns, *, kernel=True):
        self.pivots = {}
        self.kernel = []
        for index, v in enumerate(columns):
            coeff = 1 << index
            while v:
                t = v.bit_length() - 1
                held = self.pivots.get(t)
                if held is None:
                    self.pivots[t] = (v, coeff)
                    break
                v ^= held[0]
                coeff ^= held[1]
            if not v and kernel:
                self.kernel.append(coeff)

    def solve(self, value):
        coeff = 0
        while value:
            held = self.pivots.get(value.bit_length() - 1)
            if held is None:
                return None
            value ^= held[0]
            coeff ^= held[1]
        return coeff


def _combine(columns, mask):
    out = 0
    while mask:
        low = mask & -mask
        out ^= columns[low.bit_length() - 1]
        mask ^= low
    return out


def _intersection(left, right):
    """Basis of span(left) /\\ span(right) for independent inputs: the dependencies of their concatenation,
    projected on the left part (cdist/tree_dp.py:111-116)."""
    solver = _LinearMap(list(left) + list(right))
    mask = (1 << len(left)) - 1
    return tuple(_combine(left, v & mask) for v in solver.kernel)


@dataclass(frozen=True)
class TreeNode:
    leaf: int | None
    left: int | None = None
    right: int | None = None


@dataclass(frozen=True)
class TreeJoin:
    left_rank: int
    right_rank: int
    parent_rank: int
    parent_maps: tuple
    free_maps: tuple

    @property
    def join_dimension(self):
        return self.parent_rank + len(self.free_maps)

    @property
    def work(self):
        return 1 << self.join_dimension


def _tree_nodes(tree, ncols):
    """Postorder node list of the tree, validated without recursion (cdist/tree_dp.py:144-174)."""
    if ncols == 0:
        if tree is not None:
            raise ValueError("the empty problem requires tree=None")
        return ()
    nodes, stack, done, seen, visits = [], [(tree, False)], [], set(), 0
    while stack:
        item, finish = stack.pop()
        if finish:
            right, left = done.pop(), done.pop()
            done.append(len(nodes))
            nodes.append(TreeNode(None, left, right))
            continue
        visits += 1
        if visits > 2 * ncols - 1:
            raise ValueError("tree has too many nodes, repeated leaves, or a cycle")
        if isinstance(item, int) and not isinstance(item, bool):
            if not 0 <= item < ncols or item in seen:
                raise ValueError("tree leaves must be each column index exactly once")
            seen.add(item)
            done.append(len(nodes))
            nodes.append(TreeNode(item))
        elif isinstance(item, (tuple, list)) and len(item) == 2:
            stack.extend(((item, True), (item[1], False), (item[0], False)))
        else:
            raise ValueError("each tree node must be an integer leaf or a binary pair")
    if len(seen) != ncols or len(done) != 1:
        raise ValueError("tree leaves must be each column index exactly once")
    return tuple(nodes)


def _owned_size(value):
    """Recursive CPython size of an object graph, shared objects counted once."""
    total, seen, stack = 0, set(), [value]
    while stack:
        item = stack.pop()
        if id(item) in seen:
            continue
        seen.add(id(item))
        total += sys.getsizeof(item)
        if isinstance(item, dict):
            stack.extend(item.keys())
            stack.extend(item.values())
        elif isinstance(item, (tuple, list, set)):
            stack.extend(item)
        elif hasattr(item, "__dict__"):
            stack.append(vars(item))
    return total


def _cost_dtype(ncols):
    """Smallest unsigned dtype holding ncols + 1, the infinity (cdist/tree_dp.py:195-199)."""
    for dt in (np.uint8, np.uint16, np.uint32, np.uint64):
        if ncols + 1 <= np.iinfo(dt).max:
            return np.dtype(dt)
    raise ValueError("number of columns exceeds supported exact integer range")


class CompiledTreeDP:
    """A binary tree realization of M, reusable for arbitrary affine targets (cdist/tree_dp.py:235-505).

    max_compile_bytes bounds a conservative estimate of the basis storage before it is built; max_compile_seconds
    is checked between nodes.  Neither compilation nor preflight allocates an exponential-size array."""

    def __init__(self, columns, tree, *, max_compile_bytes=DEFAULT_MEMORY_BYTES, max_compile_seconds=60.0):
        start = time.monotonic()
        self.columns = tuple(columns)
        if any(not isinstance(v, int) or isinstance(v, bool) or v < 0 for v in self.columns):
            raise ValueError("columns must be nonnegative Python integers")
        self.ncols = len(self.columns)
        self.nodes = _tree_nodes(tree, self.ncols)
        self.root = len(self.nodes) - 1
        self.dtype = _cost_dtype(self.ncols).name
        self.itemsize = np.dtype(self.dtype).itemsize
        self.infinity = self.ncols + 1

        def check_time():
            if max_compile_seconds is not None and time.monotonic() - start > max_compile_seconds:
                raise TreeDPResourceError("tree compilation exceeded max_compile_seconds")

        rank_bound = min(self.ncols, max((v.bit_length() for v in self.columns), default=0))
        leaves_below = []
        for node in self.nodes:
            leaves_below.append(1 if node.leaf is not None
                                else leaves_below[node.left] + leaves_below[node.right])
        vector_bytes = sys.getsizeof(1 << rank_bound) + 8
        ambient_bytes = max((sys.getsizeof(v) for v in self.columns), default=28) + 8
        self.compile_storage_bound_bytes = (
            _owned_size(self.columns) + _owned_size(self.nodes)
            + sum(min(rank_bound, c) for c in leaves_below) * (4 * ambient_bytes + 2 * vector_bytes)
            + len(self.nodes) * rank_bound * (ambient_bytes + vector_bytes)
            + 16 * (rank_bound + 1) ** 2 + len(self.nodes) * 512 + 16384)
        if max_compile_bytes is not None and self.compile_storage_bound_bytes > max_compile_bytes:
            raise TreeDPResourceError(
                "conservative compilation bound %d bytes exceeds max_compile_bytes=%d"
                % (self.compile_storage_bound_bytes, max_compile_bytes),
                {"compile_storage_bound_bytes": self.compile_storage_bound_bytes})
        self.solver = _LinearMap(self.columns, kernel=False)
        self.matrix_rank = len(self.solver.pivots)

        # span of every subtree (bottom-up), then its complement and interface (top-down)
        spans, span_rank = {}, []
        for index, node in enumerate(self.nodes):
            check_time()
            spans[index] = (_basis((self.columns[node.leaf],)) if node.leaf is not None
                            else _basis(spans[node.left] + spans[node.right]))
            span_rank.append(len(spans[index]))
        outside = {self.root: ()} if self.nodes else {}
        interfaces = [()] * len(self.nodes)
        outside_rank = [0] * len(self.nodes)
        for index in reversed(range(len(self.nodes))):
            check_time()
            node = self.nodes[index]
            comp = outside.pop(index)
            outside_rank[index] = len(comp)
            interfaces[index] = _intersection(spans.pop(index), comp)
            if node.leaf is None:
                outside[node.left] = _basis(comp + spans[node.right])
                outside[node.right] = _basis(comp + spans[node.left])
        self.interfaces = tuple(interfaces)

        # join maps, the J identity, and the child order minimising the live arena
        self.joins, self.execution_children, peaks = {}, {}, []
        for index, node in enumerate(self.nodes):
            check_time()
            prank = len(interfaces[index])
            size = 1 << prank
            if node.leaf is not None:
                peaks.append(size)
                continue
            left, right = interfaces[node.left], interfaces[node.right]
            solver = _LinearMap(list(left) + list(right))
            maps = tuple(solver.solve(v) for v in interfaces[index])
            if any(m is None for m in maps):
                raise AssertionError("parent interface must be contained in the child sum")
            join = TreeJoin(len(left), len(right), prank, maps, tuple(solver.kernel))
            if join.join_dimension != (span_rank[node.left] + span_rank[node.right] + outside_rank[index]
                                       - self.matrix_rank):
                raise AssertionError("compatible-pair dimension identity failed")
            self.joins[index] = join
            lsize, rsize = 1 << len(left), 1 << len(right)
            local = lsize + rsize + size
            lr = max(peaks[node.left], lsize + peaks[node.right], local)
            rl = max(peaks[node.right], rsize + peaks[node.left], local)
            if lr <= rl:
                self.execution_children[index] = (node.left, node.right)
                peaks.append(lr)
            else:
                self.execution_children[index] = (node.right, node.left)
                peaks.append(rl)
        order, stack = [], ([(self.root, False)] if self.nodes else [])
        while stack:
            index, finish = stack.pop()
            if finish or self.nodes[index].leaf is not None:
                order.append(index)
            else:
                first, second = self.execution_children[index]
                stack.extend(((index, True), (second, False), (first, False)))
        self.order = tuple(order)
        self.peak_live_entries = peaks[self.root] if self.nodes else 0
        self.compilation_seconds = time.monotonic() - start
        self.compiled_metadata_bytes = _owned_size(self)

    def preflight(self, *, max_memory_bytes=DEFAULT_MEMORY_BYTES, max_join_work=DEFAULT_JOIN_WORK,
                  max_table_bits=None):
        """Exact device-arena size, candidate-pair counts and rejection reasons (cdist/tree_dp.py:349-399)."""
        ranks = [len(v) for v in self.interfaces]
        max_rank = max(ranks, default=0)
        max_join = max((j.join_dimension for j in self.joins.values()), default=0)
        work = sum(j.work for j in self.joins.values())
        table_bytes = self.peak_live_entries * self.itemsize
        scratch_bytes = 2 * max_join * 8
        peak = table_bytes + scratch_bytes
        reasons = []
        if max_rank >= 63 or max_join >= 63:
            reasons.append("table or join index requires at least 63 bits")
        if table_bytes > np.iinfo(np.intp).max:
            reasons.append("arena exceeds the platform array index range")
        if max_table_bits is not None and max_rank > max_table_bits:
            reasons.append("interface rank %d exceeds max_table_bits=%d" % (max_rank, max_table_bits))
        if max_memory_bytes is not None and peak > max_memory_bytes:
            reasons.append("dense workspace %d bytes exceeds max_memory_bytes=%d" % (peak, max_memory_bytes))
        if max_join_work is not None and work > max_join_work:
            reasons.append("join work %d exceeds max_join_work=%d" % (work, max_join_work))
        return {
            "allowed": not reasons, "rejection_reasons": reasons,
            "ncols": self.ncols, "matrix_rank": self.matrix_rank,
            "dtype": self.dtype, "dtype_bytes": self.itemsize, "infinity": self.infinity,
            "edge_state_dimensions": ranks, "max_interface_rank": max_rank,
            "max_join_dimension": max_join, "join_work": work,
            "join_profiles": [dict(node=i, left_rank=j.left_rank, right_rank=j.right_rank,
                                   parent_rank=j.parent_rank, kernel_rank=len(j.free_maps),
                                   join_dimension=j.join_dimension, work=j.work,
                                   child_parent_table_bytes=self.itemsize * (
                                       (1 << j.left_rank) + (1 << j.right_rank) + (1 << j.parent_rank)))
                              for i, j in self.joins.items()],
            "largest_table_bytes": (1 << max_rank) * self.itemsize if self.nodes else 0,
            "peak_live_table_bytes": table_bytes, "arena_bytes": table_bytes,
            "scratch_bytes": scratch_bytes, "peak_dense_bytes": peak,
            "compiled_metadata_bytes": self.compiled_metadata_bytes,
            "compile_storage_bound_bytes": self.compile_storage_bound_bytes,
            "compilation_seconds": self.compilation_seconds,
            "memory_scope": "device arena and generator maps; Python plan reported separately",
            "threads": "gpu",
        }

    def run(self, target=0, *, max_memory_bytes=DEFAULT_MEMORY_BYTES, max_join_work=DEFAULT_JOIN_WORK,
            max_table_bits=None, max_seconds=60.0, backend="hip", stream=None):
        """(exact minimum weight or None, statistics), cdist/tree_dp.py:401-505.  backend "hip" (alias "c": the
        native path) runs every join on the GPU; time is checked between launches."""
        if not isinstance(target, int) or isinstance(target, bool) or target < 0:
            raise ValueError("target must be a nonnegative Python integer")
        if backend not in ("hip", "c"):
            raise ValueError("backend must be 'hip' (or its alias 'c')")
        profile = self.preflight(max_memory_bytes=max_memory_bytes, max_join_work=max_join_work,
                                 max_table_bits=max_table_bits)
        e0 = self.solver.solve(target)
        if e0 is None:
            return None, dict(profile, feasible=False, seconds=0.0, join_work_done=0, allocated_dense_bytes=0,
                              backend="hip")
        if not profile["allowed"]:
            raise TreeDPResourceError("; ".join(profile["rejection_reasons"]), profile)
        if not self.nodes:
            return 0, dict(profile, feasible=True, seconds=0.0, join_work_done=0, allocated_dense_bytes=0,
                           backend="hip")
        L = _lib.lib()
        start = time.monotonic()

        def check_time():
            if max_seconds is not None and time.monotonic() - start > max_seconds:
                raise TreeDPResourceError("tree DP exceeded max_seconds; no exact result returned", profile)

        check_time()
        item = self.itemsize
        arena = _lib.device_empty(self.peak_live_entries * item)
        base = arena.data_ptr()
        sh = _lib.stream_handle(stream)
        maps = np.zeros((2, max(1, profile["max_join_dimension"])), dtype=np.uint64)
        stack, live, observed_peak, done = [], 0, 0, 0
        for index in self.order:
            check_time()
            node = self.nodes[index]
            size = 1 << len(self.interfaces[index])
            observed_peak = max(observed_peak, live + size)
            out = base + live * item
            if node.leaf is not None:
                bit = (e0 >> node.leaf) & 1
                v0 = bit if self.columns[node.leaf] else 0      # a zero column is free for every e0
                _lib.check(L.tdp_tree_leaf(item, out, size, v0, bit ^ 1, self.infinity, sh), "tdp_tree_leaf")
                stack.append((index, live, size))
                live += size
                continue
            second, first = stack.pop(), stack.pop()
            left, right = (first, second) if first[0] == node.left else (second, first)
            join = self.joins[index]
            gens = join.parent_maps + join.free_maps
            lmask = (1 << join.left_rank) - 1
            for b, g in enumerate(gens):
                maps[0, b] = g & lmask
                maps[1, b] = g >> join.left_rank
            _lib.check(L.tdp_tree_fill(item, out, size, self.infinity, sh), "tdp_tree_fill")
            lo, ro = base + left[1] * item, base + right[1] * item
            for begin in range(0, join.work, _CHUNK_PAIRS):
                check_time()
                end = min(begin + _CHUNK_PAIRS, join.work)
                _lib.check(L.tdp_tree_join(item, lo, ro, out, maps[0].ctypes.data, maps[1].ctypes.data,
                                           join.parent_rank, join.join_dimension, begin, end, self.infinity, sh),
                           "tdp_tree_join")
                done += end - begin
            dest = first[1]
            _lib.check(L.tdp_tree_move(base + dest * item, out, size * item, sh), "tdp_tree_move")
            live = dest + size
            stack.append((index, dest, size))
        check_time()
        if live != 1 or len(stack) != 1 or observed_peak != self.peak_live_entries:
            raise AssertionError("arena schedule disagrees with exact preflight")
        _lib.torch().cuda.synchronize()
        value = int(arena[:item].cpu().numpy().view(self.dtype)[0])
        if value == self.infinity:
            raise AssertionError("a consistent affine system must have a finite root cost")
        return value, dict(profile, feasible=True, seconds=time.monotonic() - start, join_work_done=done,
                           allocated_dense_bytes=self.peak_live_entries * item + 16 * profile["max_join_dimension"],
                           observed_peak_live_table_bytes=observed_peak * item, backend="hip")


def _augmented(h_cols, l_cols, mask):
    nrows = max((v.bit_length() for v in h_cols), default=0)
    return nrows, tuple(h | ((_popcount(l & mask) & 1) << nrows) for h, l in zip(h_cols, l_cols))


def _check_query(h_cols, l_cols, mask):
    if len(h_cols) != len(l_cols):
        raise ValueError("H and L must have the same number of columns")
    if not isinstance(mask, int) or isinstance(mask, bool) or mask < 0:
        raise ValueError("observable mask must be a nonnegative Python integer")
    if any(not isinstance(v, int) or isinstance(v, bool) or v < 0 for v in (*h_cols, *l_cols)):
        raise ValueError("columns must be nonnegative Python integers")


def tree_min_distance_for_mask(h_cols, l_cols, mask, tree, *, max_compile_bytes=DEFAULT_MEMORY_BYTES,
                               max_compile_seconds=60.0, **limits):
    """Solve H e = 0, parity(mask & L e) = 1 with the given binary tree (cdist/tree_dp.py:508-527): the logical
    parity becomes one more row, and the target is that row."""
    _check_query(h_cols, l_cols, mask)
    nrows, columns = _augmented(h_cols, l_cols, mask)
    program = CompiledTreeDP(columns, tree, max_compile_bytes=max_compile_bytes,
                             max_compile_seconds=max_compile_seconds)
    distance, info = program.run(1 << nrows, **limits)
    info["observable_mask"] = mask
    return distance, info


def preflight_tree_dp(h_cols, l_cols, mask, tree, *, max_compile_bytes=DEFAULT_MEMORY_BYTES,
                      max_compile_seconds=60.0, **limits):
    """Compile the logical query and return resource and feasibility data (cdist/tree_dp.py:530-546)."""
    _check_query(h_cols, l_cols, mask)
    nrows, columns = _augmented(h_cols, l_cols, mask)
    program = CompiledTreeDP(columns, tree, max_compile_bytes=max_compile_bytes,
                             max_compile_seconds=max_compile_seconds)
    return dict(program.preflight(**limits), feasible=program.solver.solve(1 << nrows) is not None,
                observable_mask=mask)
