"""GPU parity of the tree-realization DP: joins, leaves and arena compaction on the device against brute force,
the trellis DP and the oracle's CPU run of the same compiled program."""
import random
from itertools import product
from unittest.mock import patch

import pytest

import oracle
from test_oracle_tree import balanced, comb, random_tree
from tropidist_amd import kernel, problem, trellis, tree_dp
from tropidist_amd.tree_dp import CompiledTreeDP, tree_min_distance_for_mask

pytestmark = pytest.mark.gpu


def test_all_small_matrices_all_targets():
    print("case tree 2728 queries", flush=True)
    solved = 0
    for ncols in range(5):
        tree = balanced(list(range(ncols)))
        for columns in product(range(4), repeat=ncols):
            ref = oracle.brute_affine(columns)
            program = CompiledTreeDP(columns, tree)
            for target in range(8):
                got, stats = program.run(target)
                assert got == ref.get(target), (columns, target)
                if stats["feasible"] and ncols:
                    assert stats["allocated_dense_bytes"] == stats["peak_dense_bytes"]
                    assert stats["observed_peak_live_table_bytes"] == stats["arena_bytes"]
                    assert stats["join_work_done"] == stats["join_work"]
                solved += 1
    assert solved == 2728


def test_random_logical_queries_against_trellis():
    print("case tree random vs trellis", flush=True)
    rng = random.Random(20261021)
    for trial in range(160):
        ncols, nrows, nobs = rng.randrange(1, 13), rng.randrange(0, 7), rng.randrange(1, 4)
        h = [rng.randrange(1 << nrows) for _ in range(ncols)]
        l = [rng.randrange(1 << nobs) for _ in range(ncols)]
        order = list(range(ncols))
        rng.shuffle(order)
        mask = rng.randrange(1 << nobs)
        want = oracle.brute(h, l, mask)
        assert kernel.run_plan(trellis.build_plan(h, l, mask, order, nrows))[0] == want
        tree = (comb(order), balanced(order), random_tree(order, rng))[trial % 3]
        assert tree_min_distance_for_mask(h, l, mask, tree)[0] == want, trial


def test_large_distance_selects_uint16():
    print("case tree uint16 chain", flush=True)
    for ncols in (254, 255, 300):
        h = [1] + [(1 << (j - 1)) | (1 << j) for j in range(1, ncols - 1)] + [1 << (ncols - 2)]
        d, stats = tree_min_distance_for_mask(h, [1] + [0] * (ncols - 1), 1, balanced(list(range(ncols))))
        assert d == ncols and stats["dtype"] == ("uint8" if ncols == 254 else "uint16")


def test_chunked_joins_on_gpu():
    print("case tree chunks of 3", flush=True)
    columns = [1, 2, 4, 7, 3, 5, 6, 1]
    with patch.object(tree_dp, "_CHUNK_PAIRS", 3):
        program = CompiledTreeDP(columns, ((0, (1, (2, 3))), ((4, 5), (6, 7))))
        for target, want in oracle.brute_affine(columns).items():
            assert program.run(target)[0] == want


def test_overlapping_compaction_and_deep_comb():
    print("case tree compaction / comb", flush=True)
    program = CompiledTreeDP([1 << j for j in range(8)] * 2, balanced(list(range(16))))
    for target in (0, 1, 3, 17, 127, 128, 255):
        d, stats = program.run(target)
        assert d == bin(target).count("1") and stats["observed_peak_live_table_bytes"] == 544
    assert CompiledTreeDP([0] * 1300, comb(list(range(1300)))).run()[0] == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_banded_balanced_tree_equals_oracle(seed):
    """Interface ranks up to ~14 with wide joins: the GPU run equals the oracle's CPU run and the trellis DP."""
    print("case tree banded seed", seed, flush=True)
    p = problem.banded(R=40, N=160, kmin=2, kmax=4, W=12, nobs=1, rho=0.2, seed=seed)
    order = list(range(p.ncols))
    plan = trellis.build_plan(p.h, p.l, 1, order, p.nrows)
    nrows = p.nrows
    aug = [h | ((l & 1) << nrows) for h, l in zip(p.h, p.l)]
    program = CompiledTreeDP(aug, balanced(order))
    prof = program.preflight(max_memory_bytes=None, max_join_work=None)
    got, _ = program.run(1 << nrows, max_memory_bytes=None, max_join_work=None)
    assert got == oracle.tree_run(program, 1 << nrows) == kernel.run_plan(plan)[0]
    assert prof["max_join_dimension"] >= 8


@pytest.mark.parametrize("itemsize", [1, 2, 4, 8])
def test_typed_join_entry_points_equal_oracle(itemsize):
    """tdp_tree_join_u8/16/32/64 (the reference's tree_join_uX names) on random tables and generator maps, whole
    and in Gray-counter ranges, against the oracle's join of the same item size; inf sits near the dtype's top so
    the overflow guard (a + b <= inf) is exercised."""
    import numpy as np
    from tropidist_amd import _lib
    print("case tree typed join u%d" % (8 * itemsize), flush=True)
    t = _lib.torch()
    L = _lib.lib()
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[itemsize]
    inf = int(np.iinfo(dt).max) - 1
    rng = np.random.default_rng(40 + itemsize)
    for trial in range(12):
        lr, rr = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        pbits = int(rng.integers(0, min(lr, rr) + 1))
        gens = pbits + int(rng.integers(0, 4))
        lm = rng.integers(0, 1 << lr, size=max(1, gens), dtype=np.uint64)
        rm = rng.integers(0, 1 << rr, size=max(1, gens), dtype=np.uint64)
        left = rng.integers(0, inf + 1, size=1 << lr, dtype=np.uint64).astype(dt)
        right = rng.integers(0, inf + 1, size=1 << rr, dtype=np.uint64).astype(dt)
        if trial % 2:       # small costs, so most sums stay below inf
            left = (left % 50).astype(dt)
            right = (right % 50).astype(dt)
        total = 1 << gens
        cuts = sorted({0, total, int(rng.integers(0, total + 1)), int(rng.integers(0, total + 1))})
        want = np.full(1 << pbits, inf, dtype=dt)
        for a, b in zip(cuts, cuts[1:]):
            oracle.tree_join(itemsize, left, right, want, lm, rm, pbits, gens, a, b, inf)
        dl, dr = (t.from_numpy(x.view(np.uint8).copy()).cuda() for x in (left, right))
        do = t.from_numpy(np.full(1 << pbits, inf, dtype=dt).view(np.uint8).copy()).cuda()
        f = getattr(L, "tdp_tree_join_u%d" % (8 * itemsize))
        for a, b in zip(cuts, cuts[1:]):
            _lib.check(f(dl.data_ptr(), dr.data_ptr(), do.data_ptr(), lm.ctypes.data, rm.ctypes.data, pbits, gens,
                         a, b, inf, _lib.stream_handle()), "tdp_tree_join_u%d" % (8 * itemsize))
        t.cuda.synchronize()
        got = do.cpu().numpy().view(dt)
        assert (got == want).all(), (itemsize, trial)


@pytest.mark.parametrize("itemsize", [1, 2])
def test_tiled_full_joins_equal_oracle(itemsize):
    """Full joins that take the LDS-tiled kernel (parent ranks 12-16, 4-10 kernel generators, parent directions
    that shift the kernel index on one or both sides) against the oracle's Gray-code join, with the infinity in the
    tables; the plan must be the tiled one for most of them."""
    import numpy as np
    from test_tree_tiles import oracle_rank, plan, xor_of
    from tropidist_amd import _lib
    print("case tree tiled joins u%d" % (8 * itemsize), flush=True)
    t = _lib.torch()
    L = _lib.lib()
    dt = np.uint8 if itemsize == 1 else np.uint16
    rng = np.random.default_rng(70 + itemsize)
    tiled = 0
    for trial in range(16):
        kr = int(rng.integers(4, 11))
        pr = int(rng.integers(12, 17))
        lr, rr = kr + int(rng.integers(0, 5)), kr + int(rng.integers(0, 5))
        while True:
            A = [int(x) for x in rng.integers(1, 1 << lr, size=kr)]
            B = [int(x) for x in rng.integers(1, 1 << rr, size=kr)]
            if oracle_rank(A) == kr and oracle_rank(B) == kr:
                break
        lm = [int(x) for x in rng.integers(0, 1 << lr, size=pr)] + A
        rm = [int(x) for x in rng.integers(0, 1 << rr, size=pr)] + B
        for i in range(pr):
            r = rng.random()
            if r < 0.4:
                lm[i] = xor_of(A, int(rng.integers(0, 1 << kr)))
            if r > 0.6 or r < 0.2:
                rm[i] = xor_of(B, int(rng.integers(0, 1 << kr)))
        J = pr + kr
        inf = int(rng.choice([60, 200 if itemsize == 1 else 6001]))
        tiled += plan(lm, rm, pr, J, itemsize, inf, min_tiles_log=10)[1] > 0
        left = np.where(rng.random(1 << lr) < 0.1, inf, rng.integers(0, inf, size=1 << lr)).astype(dt)
        right = np.where(rng.random(1 << rr) < 0.1, inf, rng.integers(0, inf, size=1 << rr)).astype(dt)
        lmv, rmv = np.array(lm, dtype=np.uint64), np.array(rm, dtype=np.uint64)
        want = np.full(1 << pr, inf, dtype=dt)
        oracle.tree_join(itemsize, left, right, want, lmv, rmv, pr, J, 0, 1 << J, inf)
        dl, dr = (t.from_numpy(x.view(np.uint8).copy()).cuda() for x in (left, right))
        do = t.from_numpy(np.full(1 << pr, inf, dtype=dt).view(np.uint8).copy()).cuda()
        _lib.check(L.tdp_tree_join(itemsize, dl.data_ptr(), dr.data_ptr(), do.data_ptr(), lmv.ctypes.data,
                                   rmv.ctypes.data, pr, J, 0, 1 << J, inf, _lib.stream_handle()), "tdp_tree_join")
        t.cuda.synchronize()
        got = do.cpu().numpy().view(dt)
        assert (got == want).all(), (itemsize, trial, int((got != want).sum()))
    assert tiled >= 8, tiled


def test_bench_tree_instance_with_tiled_joins():
    """The tree instance of the benchmark (banded R=75 N=300 W=10, balanced tree): its widest joins take the tiled
    kernel, and the distance equals the trellis DP's and the oracle's CPU run of the same program."""
    print("case tree bench instance tiled", flush=True)
    p = problem.banded(R=75, N=300, kmin=2, kmax=5, W=10, nobs=1, rho=0.1, seed=10)
    aug = [h | ((l & 1) << p.nrows) for h, l in zip(p.h, p.l)]
    program = CompiledTreeDP(aug, balanced(list(range(p.ncols))))
    from test_tree_tiles import plan
    ntiled = sum(plan(list(j.lmaps), list(j.rmaps), j.parent_rank, j.join_dimension, program.itemsize,
                      program.infinity, min_tiles_log=10)[1] > 0 for j in program.joins.values() if j.lmaps)
    assert ntiled >= 10, ntiled
    lim = dict(max_join_work=1 << 40, max_memory_bytes=1 << 36, max_seconds=600.0)
    got, _ = program.run(1 << p.nrows, **lim)
    dense = kernel.run_plan(trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows))[0]
    assert got == dense == oracle.tree_run(program, 1 << p.nrows)
