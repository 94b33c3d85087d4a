"""Entry points of the C-ABI that the Python layers do not call themselves, each checked against what pins it:
tdp_update against the oracle's dp_update (cdist/viterbi.c:72-118), tdp_tree_leaf / tdp_tree_fill /
tdp_tree_move against their closed forms (the numpy fills of cdist/tree_dp.py:446-454 and the memmove of
cdist/tree_dp.c:46-51), and the device search's tuning calls (tdp_gm_set_chunk, tdp_gm_last_launches,
tdp_gm_release) against the search's own results with their defaults."""

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures
from tropidist_amd import _lib, search
from tropidist_amd.problem import Problem

pytestmark = pytest.mark.gpu


def _to_device(a):
    t = _lib.torch()
    return t.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("bits", [1, 2, 7, 12, 20])
def test_update_equals_oracle(bits):
    """T[s] <- min(T[s], T[s ^ h] + 1), saturating at 255, in place, for several h including the top bit."""
    print("case tdp_update bits", bits, flush=True)
    L = _lib.lib()
    rng = np.random.default_rng(100 + bits)
    n = 1 << bits
    for h in sorted({1, n - 1, n >> 1, int(rng.integers(1, n)) if n > 1 else 1}):
        T = rng.integers(0, 256, size=n, dtype=np.uint8)
        T[rng.integers(0, n, size=max(1, n // 8))] = 255
        want = T.copy()
        oracle.olib().or_dp_update(want.ctypes.data, n, h)
        d = _to_device(T)
        _lib.check(L.tdp_update(d.data_ptr(), n, h, _lib.stream_handle()), "tdp_update")
        assert np.array_equal(d.cpu().numpy(), want), (bits, h)


@pytest.mark.parametrize("itemsize", [1, 2, 4, 8])
def test_tree_leaf_and_fill(itemsize):
    """Leaf tables hold v0, v1 and then the fill value; fill sets every item; sizes 1, 2 and 300,001."""
    print("case tree leaf/fill itemsize", itemsize, flush=True)
    L = _lib.lib()
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[itemsize]
    top = (1 << (8 * itemsize)) - 1
    for size in (1, 2, 300001):
        d = _to_device(np.zeros(size * itemsize, dtype=np.uint8))
        v0, v1, fill = 0, min(3, top), top
        _lib.check(L.tdp_tree_leaf(itemsize, d.data_ptr(), size, v0, v1, fill, _lib.stream_handle()),
                   "tdp_tree_leaf")
        want = np.full(size, fill, dtype=dt)
        want[0] = v0
        if size > 1:
            want[1] = v1
        assert np.array_equal(d.cpu().numpy().view(dt), want), size
        _lib.check(L.tdp_tree_fill(itemsize, d.data_ptr(), size, top - 1, _lib.stream_handle()), "tdp_tree_fill")
        assert np.array_equal(d.cpu().numpy().view(dt), np.full(size, top - 1, dtype=dt)), size
    assert L.tdp_tree_leaf(3, d.data_ptr(), 4, 0, 0, 0, _lib.stream_handle()) == _lib.TDP_EINVAL
    assert L.tdp_tree_leaf(itemsize, d.data_ptr(), 0, 0, 0, 0, _lib.stream_handle()) == _lib.TDP_EINVAL


@pytest.mark.parametrize("nbytes,gap", [(1000, 1), (1000, 999), (500000, 4096), (3 << 20, (1 << 20) + 7),
                                        (2 << 20, 3 << 20), (70000, 70000)])
def test_tree_move_is_memmove_toward_lower_addresses(nbytes, gap):
    """Overlapping and disjoint moves equal memmove; a move toward a higher address is refused."""
    print("case tree move", nbytes, gap, flush=True)
    L = _lib.lib()
    rng = np.random.default_rng(nbytes + gap)
    buf = rng.integers(0, 256, size=nbytes + gap + 64, dtype=np.uint8)
    want = buf.copy()
    want[5:5 + nbytes] = buf[5 + gap:5 + gap + nbytes].copy()
    d = _to_device(buf)
    base = d.data_ptr()
    _lib.check(L.tdp_tree_move(base + 5, base + 5 + gap, nbytes, _lib.stream_handle()), "tdp_tree_move")
    assert np.array_equal(d.cpu().numpy(), want)
    assert L.tdp_tree_move(base + 6, base + 5, 16, _lib.stream_handle()) == _lib.TDP_EINVAL


def _code(name):
    for r in dense_fixtures():
        if r["name"] == name:
            return (Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"]),
                    min(x for x in r["distances"] if x is not None))
    pytest.skip("fixture %s missing" % name)


def test_search_chunk_launches_and_release():
    """A smaller node chunk gives the same distance, solutions and node counts in more launches; after
    tdp_gm_release frees the kept buffers the next search allocates them again and gives the same result."""
    print("case search chunk/launches/release", flush=True)
    L = _lib.lib()
    p, d = _code("code_bb90")
    base = search.distance(p, k_table=2, kmax=d, enum_all=True, time_limit=120.0)
    launches = L.tdp_gm_last_launches()
    assert base["distance"] == d and launches >= 1
    try:
        L.tdp_gm_set_chunk(256)                     # 256 nodes a launch (the floor is one 128-thread block)
        small = search.distance(p, k_table=2, kmax=d, enum_all=True, time_limit=120.0)
        small_launches = L.tdp_gm_last_launches()
    finally:
        L.tdp_gm_set_chunk(0)
    widest = max(base["runs"][-1]["nodes"])
    assert small_launches >= launches
    if 256 < widest <= 1 << 19:                     # a depth the default chunk expands in one launch
        assert small_launches > launches, (small_launches, launches, widest)
    assert small["distance"] == base["distance"] and small["solutions"] == base["solutions"]
    assert [r["nodes"] for r in small["runs"]] == [r["nodes"] for r in base["runs"]]
    L.tdp_gm_release()
    again = search.distance(p, k_table=2, kmax=d, enum_all=True, time_limit=120.0)
    assert again["distance"] == base["distance"] and again["solutions"] == base["solutions"]
    assert L.tdp_gm_last_launches() == launches
