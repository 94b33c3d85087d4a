"""Tree-realization DP: the host compilation (interfaces, join maps, J identity, preflight arithmetic, resource
rejections) and the oracle's joins, against brute force and the reference's compiled tree_dp.c."""
import random
from itertools import product
from unittest.mock import patch

import numpy as np
import pytest

import oracle
from tropidist_amd import trellis
from tropidist_amd.tree_dp import CompiledTreeDP, TreeDPResourceError, preflight_tree_dp

needs_ref = pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")


def balanced(leaves):
    if not leaves:
        return None
    if len(leaves) == 1:
        return leaves[0]
    m = len(leaves) // 2
    return balanced(leaves[:m]), balanced(leaves[m:])


def comb(leaves):
    if not leaves:
        return None
    t = leaves[0]
    for x in leaves[1:]:
        t = t, x
    return t


def random_tree(leaves, rng):
    if len(leaves) == 1:
        return leaves[0]
    s = rng.randrange(1, len(leaves))
    return random_tree(leaves[:s], rng), random_tree(leaves[s:], rng)


def test_all_small_matrices_all_targets_oracle():
    """tests/test_tree_dp.py:72-89: all 2-row matrices with <= 4 columns, all 8 targets: 2728 queries."""
    solved = 0
    for ncols in range(5):
        tree = balanced(list(range(ncols)))
        for columns in product(range(4), repeat=ncols):
            ref = oracle.brute_affine(columns)
            program = CompiledTreeDP(columns, tree)
            for target in range(8):
                assert oracle.tree_run(program, target) == ref.get(target), (columns, target)
                solved += 1
    assert solved == 2728


def test_random_queries_interfaces_and_join_identity():
    """tests/test_tree_dp.py:91-123: interface ranks equal cut ranks; J equals the rank identity; the oracle run
    equals brute force."""
    rng = random.Random(20261021)
    for trial in range(160):
        ncols, nrows, nobs = rng.randrange(1, 13), rng.randrange(0, 7), rng.randrange(1, 4)
        h = [rng.randrange(1 << nrows) for _ in range(ncols)]
        l = [rng.randrange(1 << nobs) for _ in range(ncols)]
        order = list(range(ncols))
        rng.shuffle(order)
        mask = rng.randrange(1 << nobs)
        want = oracle.brute(h, l, mask)
        tree = (comb(order), balanced(order), random_tree(order, rng))[trial % 3]
        aug = [h[j] | ((bin(l[j] & mask).count("1") & 1) << nrows) for j in range(ncols)]
        program = CompiledTreeDP(aug, tree)
        assert oracle.tree_run(program, 1 << nrows) == want, trial
        supports = []
        for index, node in enumerate(program.nodes):
            sup = {node.leaf} if node.leaf is not None else supports[node.left] | supports[node.right]
            supports.append(sup)
            before = list(sup) + [j for j in range(ncols) if j not in sup]
            assert len(program.interfaces[index]) == trellis.cut_profile(aug, before)[len(sup)]
            if node.leaf is None:
                rank = lambda s: trellis.rank_profile(aug, list(s))[-1]
                expected = rank(supports[node.left]) + rank(supports[node.right]) \
                    + rank(set(range(ncols)) - sup) - rank(range(ncols))
                assert program.joins[index].join_dimension == expected


def test_preflight_arithmetic_and_dtype():
    """tests/test_tree_dp.py:133-141, 196-208: uint8 up to 254 columns, uint16 at 255 and 300, infinity N + 1;
    the 16 + 16-column half-identity case has 256-byte tables and a 544-byte arena."""
    for ncols, dt in ((254, "uint8"), (255, "uint16"), (300, "uint16")):
        h = [1] + [(1 << (j - 1)) | (1 << j) for j in range(1, ncols - 1)] + [1 << (ncols - 2)]
        prof = preflight_tree_dp(h, [1] + [0] * (ncols - 1), 1, balanced(list(range(ncols))))
        assert prof["dtype"] == dt and prof["infinity"] == ncols + 1 and prof["feasible"]
    program = CompiledTreeDP([1 << j for j in range(8)] * 2, balanced(list(range(16))))
    prof = program.preflight()
    assert prof["largest_table_bytes"] == 256 and prof["peak_live_table_bytes"] == 544
    assert prof["scratch_bytes"] == 16 * prof["max_join_dimension"]


def test_rejections_happen_before_any_device_allocation():
    """tests/test_tree_dp.py:149-163 with the device allocator patched to fail."""
    program = CompiledTreeDP([1, 2, 3, 1, 2, 3], ((0, (1, 2)), (3, (4, 5))))
    prof = program.preflight()
    assert prof["join_work"] > 1
    with patch("tropidist_amd._lib.device_empty", side_effect=AssertionError("allocated")):
        for kw in (dict(max_memory_bytes=prof["peak_dense_bytes"] - 1), dict(max_join_work=prof["join_work"] - 1),
                   dict(max_table_bits=0), dict(max_seconds=-1)):
            with pytest.raises(TreeDPResourceError):
                program.run(**kw)
    with pytest.raises(TreeDPResourceError):
        CompiledTreeDP([1, 2], (0, 1), max_compile_bytes=1)
    assert not preflight_tree_dp([1], [0], 1, 0)["feasible"]


def test_empty_and_malformed_inputs():
    program = CompiledTreeDP([], None)
    assert program.run(0)[0] == 0 and program.run(1)[0] is None
    for tree in ((0, 0), (0, 2), (0,), (False, 1), 0, None):
        with pytest.raises(ValueError):
            CompiledTreeDP([1, 2], tree)
    cycle = [0, None]
    cycle[1] = cycle
    with pytest.raises(ValueError):
        CompiledTreeDP([1, 2], cycle)
    for column in (-1, True, 1.0):
        with pytest.raises(ValueError):
            CompiledTreeDP([column], 0)
    for target in (-1, True, 1.0):
        with pytest.raises(ValueError):
            CompiledTreeDP([1], 0).run(target)


def test_deep_comb_compiles_without_recursion():
    program = CompiledTreeDP([0] * 1300, comb(list(range(1300))))
    assert oracle.tree_run(program) == 0 and program.preflight()["max_interface_rank"] == 0


def test_chunked_joins_restart_the_gray_code():
    """tests/test_tree_dp.py:183-194: chunks of 3 pairs."""
    columns = [1, 2, 4, 7, 3, 5, 6, 1]
    program = CompiledTreeDP(columns, ((0, (1, (2, 3))), ((4, 5), (6, 7))))
    for target, want in oracle.brute_affine(columns).items():
        assert oracle.tree_run(program, target, chunk=3) == want


@needs_ref
@pytest.mark.parametrize("item", [1, 2, 4, 8])
def test_oracle_join_equals_reference_c(item):
    rng = np.random.default_rng(item)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[item]
    for trial in range(40):
        lr, rr, pr = rng.integers(0, 7, size=3)
        kr = rng.integers(0, 5)
        J = int(pr + kr)
        inf = int(rng.integers(5, 250))
        L = rng.integers(0, inf + 1, size=1 << lr).astype(dt)
        R = rng.integers(0, inf + 1, size=1 << rr).astype(dt)
        lm = rng.integers(0, 1 << lr, size=max(1, J)).astype(np.uint64)
        rm = rng.integers(0, 1 << rr, size=max(1, J)).astype(np.uint64)
        out0 = rng.integers(0, inf + 1, size=1 << pr).astype(dt)
        for start, stop in ((0, 1 << J), (0, max(1, (1 << J) // 3)), ((1 << J) // 2, 1 << J)):
            a, b = out0.copy(), out0.copy()
            oracle.tree_join(item, L, R, a, lm, rm, int(pr), J, start, stop, inf, "oracle")
            oracle.tree_join(item, L, R, b, lm, rm, int(pr), J, start, stop, inf, "ref")
            assert np.array_equal(a, b), (trial, start, stop)
