"""The allocation-free table build (tdp_tab_build_noalloc, static device scratch) against tdp_tab_build through the
C-ABI: the same key count and the same answers to a batch of probes, on two calls in a row (the second smaller, so
stale scratch contents from the first call would show), for compact and wide (> 16 observables) tables."""
import ctypes
import random

import numpy as np
import pytest

from tropidist_amd import _lib, problem, tables

pytestmark = pytest.mark.gpu


def _build(L, p, k, log2slots, wide, zh, zl, lg, entry):
    h = L.tdp_tab_create(log2slots, 1 if wide else 0)
    assert h, L.tdp_last_error()
    rc = getattr(L, entry)(h, p.ncols, zh.ctypes.data, zl.ctypes.data, lg.ctypes.data, k)
    assert rc == 0, (entry, rc, L.tdp_last_error())
    return h


def _probe(L, h, qs):
    t = _lib.torch()
    m = len(qs)
    arrs = [np.array([q[i] for q in qs], dtype=np.uint64) for i in range(3)] + [
        np.array([q[3] for q in qs], dtype=np.int32)]
    dev = [t.from_numpy(a.view(np.int64) if a.dtype == np.uint64 else a).cuda() for a in arrs]
    out = t.empty(m, dtype=t.int32, device="cuda")
    _lib.check(L.tdp_tab_lookup_batch(h, dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), dev[3].data_ptr(),
                                      m, out.data_ptr(), _lib.stream_handle()), "tdp_tab_lookup_batch")
    t.cuda.synchronize()
    return out.cpu().numpy().tolist()


def _queries(zh, zl, lg, k, rng, n=512):
    qs = []
    ncols = len(zh)
    for _ in range(n):
        sup = rng.sample(range(ncols), rng.randint(0, min(k + 1, ncols)))
        a = b = g = 0
        for j in sup:
            a ^= int(zh[j])
            b ^= int(zl[j])
            g ^= int(lg[j])
        qs.append((a, b, g ^ rng.choice([0, 1, 2]), rng.randint(0, k)))
    return qs


@pytest.mark.parametrize("nobs", [4, 20])
def test_noalloc_build_equals_allocating_build(nobs):
    L = _lib.lib()
    rng = random.Random(1234 + nobs)
    for (R, N, k) in ((90, 400, 3), (40, 120, 3)):          # the second call is smaller
        print("case no-alloc tables nobs", nobs, "R", R, "N", N, "k", k, flush=True)
        p = problem.banded(R=R, N=N, kmin=2, kmax=5, W=12, nobs=nobs, rho=0.3, seed=R + N)
        wide = nobs > 16
        rh = tables.row_hashes(p.nrows, 777)
        zh, zl = tables.column_hashes(p.h, rh)
        lg = np.array(p.l, dtype=np.uint64)
        log2 = tables.table_log2slots(p.ncols, k)
        a = _build(L, p, k, log2, wide, zh, zl, lg, "tdp_tab_build")
        b = _build(L, p, k, log2, wide, zh, zl, lg, "tdp_tab_build_noalloc")
        try:
            ca, cb = L.tdp_tab_count(a), L.tdp_tab_count(b)
            assert ca == cb > p.ncols, (ca, cb)
            qs = _queries(zh, zl, lg, k, rng)
            want = _probe(L, a, qs)
            assert _probe(L, b, qs) == want
            assert any(x >= 0 for x in want)
        finally:
            L.tdp_tab_free(a)
            L.tdp_tab_free(b)


def test_noalloc_build_rejects_oversized_inputs():
    L = _lib.lib()
    h = L.tdp_tab_create(12, 0)
    assert h
    try:
        cols = np.zeros(4097, dtype=np.uint64)
        rc = L.tdp_tab_build_noalloc(h, 4097, cols.ctypes.data, cols.ctypes.data, cols.ctypes.data, 1)
        assert rc == _lib.TDP_EINVAL
        assert b"static scratch" in L.tdp_last_error()
    finally:
        L.tdp_tab_free(h)
