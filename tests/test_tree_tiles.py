"""The tiled form of full tree joins (tdp_tree_tile_plan, run by k_tree_join_tiled): the host plan is executed here
in numpy exactly as the kernel reads it (gathered rows, per-parent kernel shift, parent address) and must equal the
oracle's Gray-code join (cdist/tree_dp.c:13-39) on every parent, for joins of compiled banded trees and random
joins, with random tables that include the infinity."""
import ctypes
import random

import numpy as np
import pytest

import oracle
from test_oracle_tree import balanced
from tropidist_amd import _lib, problem
from tropidist_amd.tree_dp import CompiledTreeDP

MAXKR, MAXROWS, MAXOUT, MAXT = 14, 14, 40, 40


class TileArgs(ctypes.Structure):
    _fields_ = [("vl", ctypes.c_uint64 * (MAXKR + MAXROWS)), ("vr", ctypes.c_uint64 * (MAXKR + MAXROWS)),
                ("tl", ctypes.c_uint64 * MAXT), ("tr", ctypes.c_uint64 * MAXT), ("tp", ctypes.c_uint64 * MAXT),
                ("op", ctypes.c_uint64 * MAXOUT), ("od", ctypes.c_uint32 * MAXOUT)] + [
        (k, ctypes.c_int32) for k in ("kr", "s", "a", "b", "t", "lanes_log")]


def plan(lm, rm, pr, J, item, inf, min_tiles_log=0):
    L = _lib.lib()
    assert L.tdp_tree_tile_args_bytes() == ctypes.sizeof(TileArgs)
    a = np.array(lm or [0], dtype=np.uint64)
    b = np.array(rm or [0], dtype=np.uint64)
    T = TileArgs()
    tiles = ctypes.c_uint64(0)
    _lib.check(L.tdp_tree_tile_plan(a.ctypes.data, b.ctypes.data, pr, J, item, inf, min_tiles_log, ctypes.byref(T),
                                    ctypes.byref(tiles)), "tdp_tree_tile_plan")
    return T, tiles.value


def xor_of(vecs, x):
    v = 0
    k = 0
    while x:
        if x & 1:
            v ^= int(vecs[k])
        x >>= 1
        k += 1
    return v


def run_tiles(T, tiles, Ltab, Rtab, out):
    """The tiled join of the kernel, statement by statement, on numpy tables (out updated in place)."""
    kr, s, a, b = T.kr, T.s, T.a, T.b
    nod = s + a + b
    u = np.arange(1 << kr)
    addr_u_l = np.array([xor_of(T.vl, int(x)) for x in u], dtype=np.int64)
    addr_u_r = np.array([xor_of(T.vr, int(x)) for x in u], dtype=np.int64)
    for tau in range(tiles):
        lb, rb, pb = xor_of(T.tl, tau), xor_of(T.tr, tau), xor_of(T.tp, tau)
        Lt = [Ltab[lb ^ xor_of(T.vl[kr:], al) ^ addr_u_l].astype(np.int64) for al in range(1 << a)]
        Rt = [Rtab[rb ^ xor_of(T.vr[kr:], be) ^ addr_u_r].astype(np.int64) for be in range(1 << b)]
        for o in range(1 << nod):
            delta, p = xor_of(T.od, o), pb ^ xor_of(T.op, o)
            al, be = (o >> s) & ((1 << a) - 1), o >> (s + a)
            m = int((Lt[al] + Rt[be][u ^ delta]).min())
            if m < out[p]:
                out[p] = m


def check_join(lm, rm, pr, J, lr, rr, item, inf, rng, need_tiled=False):
    T, tiles = plan(list(lm), list(rm), pr, J, item, inf)
    if not tiles:
        assert not need_tiled
        return False
    dt = np.uint8 if item == 1 else np.uint16
    for _ in range(2):
        Ltab = np.array([rng.choice([inf, rng.randrange(inf + 1)]) for _ in range(1 << lr)], dtype=dt)
        Rtab = np.array([rng.choice([inf, rng.randrange(inf + 1)]) for _ in range(1 << rr)], dtype=dt)
        want = np.full(1 << pr, inf, dtype=dt)
        oracle.tree_join(item, Ltab, Rtab, want, np.array(lm, dtype=np.uint64), np.array(rm, dtype=np.uint64), pr, J,
                         0, 1 << J, inf)
        got = np.full(1 << pr, inf, dtype=np.int64)
        run_tiles(T, tiles, Ltab, Rtab, got)
        assert (got == want.astype(np.int64)).all()
    assert (1 << (T.s + T.a + T.b)) * tiles == 1 << pr
    return True


@pytest.mark.parametrize("seed", [3, 11])
def test_tiled_plans_of_banded_tree_joins_equal_oracle(seed):
    p = problem.banded(R=30, N=80, kmin=2, kmax=4, W=6, nobs=1, rho=0.2, seed=seed)
    cols = [h | ((l & 1) << p.nrows) for h, l in zip(p.h, p.l)]
    prog = CompiledTreeDP(cols, balanced(list(range(len(cols)))))
    rng = random.Random(seed)
    tiled = 0
    for node, j in sorted(prog.joins.items()):
        if j.join_dimension > 16 or not j.lmaps:
            continue
        for item, inf in ((1, 81), (2, 81)):
            tiled += check_join(j.lmaps, j.rmaps, j.parent_rank, j.join_dimension, j.left_rank, j.right_rank, item,
                                inf, rng)
    assert tiled >= 4, tiled


def test_tiled_plans_of_random_joins_equal_oracle():
    """Random generator maps: injective random kernel maps, parent maps with chosen shift / row structure."""
    rng = random.Random(7)
    done = 0
    for trial in range(60):
        kr = rng.randrange(4, 9)
        pr = rng.randrange(2, 9)
        lr, rr = kr + rng.randrange(0, 4), kr + rng.randrange(0, 4)
        while True:
            A = [rng.randrange(1, 1 << lr) for _ in range(kr)]
            B = [rng.randrange(1, 1 << rr) for _ in range(kr)]
            if oracle_rank(A) == kr and oracle_rank(B) == kr:
                break
        lm = [rng.randrange(1 << lr) for _ in range(pr)] + A
        rm = [rng.randrange(1 << rr) for _ in range(pr)] + B
        # some parent directions only shift the kernel index on one or both sides
        for i in range(pr):
            r = rng.random()
            if r < 0.3:
                lm[i] = xor_of(A, rng.randrange(1 << kr))
            if r > 0.7 or r < 0.15:
                rm[i] = xor_of(B, rng.randrange(1 << kr))
        done += check_join(lm, rm, pr, pr + kr, lr, rr, rng.choice([1, 2]), rng.choice([9, 60, 200]), rng)
    assert done >= 20, done


def test_plans_refused_outside_the_tiled_form():
    rng = random.Random(5)
    A = [1 << i for i in range(5)]
    # a kernel map that is not injective on the left
    T, tiles = plan([3, 5] + A[:4] + [A[0] ^ A[1]], [1, 2] + A[:4] + [1 << 5], 2, 7, 1, 50)
    assert tiles == 0
    # 4-byte items, and 16-bit items whose sums could overflow
    assert plan([0, 0] + A, [0, 0] + A, 2, 7, 4, 50)[1] == 0
    assert plan([0, 0] + A, [0, 0] + A, 2, 7, 2, 40000)[1] == 0
    # too few kernel generators
    assert plan([0, 0, 1, 2, 4], [0, 0, 1, 2, 4], 2, 5, 1, 50)[1] == 0
    del rng


def oracle_rank(vecs):
    basis = []
    for v in vecs:
        for b in basis:
            v = min(v, v ^ b)
        if v:
            basis.append(v)
    return len(basis)
