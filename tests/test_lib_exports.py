"""The C-ABI library loads without a GPU and exports every symbol that include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

from tropidist_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(tdp_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build libtdp.so first (python -m tropidist_amd.build)"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    """Every declared entry point has a ctypes signature in the Python binding."""
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_version_string_without_a_device():
    assert b"gfx950" in _lib.lib().tdp_version()


def test_multi_program_records_and_validation_without_a_device():
    """Record layout; tdp_run_programs_workspace sizes the workspace from counts only; tdp_run_programs validates
    the whole schedule on the host first, so a descriptor addressing outside its table is refused before any device
    work."""
    import ctypes
    import numpy as np
    from tropidist_amd import _lib, fused, problem, trellis
    L = _lib.lib()
    TdpRun = fused._run_record_type()
    assert L.tdp_run_record_bytes() == ctypes.sizeof(TdpRun) == 120
    p = problem.banded(R=40, N=200, kmin=2, kmax=5, W=8, nobs=1, rho=0.1, seed=5)
    plan = trellis.build_plan(p.h, p.l, 1, list(range(p.ncols)), p.nrows)
    segs, arr = fused.compile_native(plan, min_bits=4)
    ops = np.ascontiguousarray(plan.encode())
    runs = (TdpRun * 1)()
    r = runs[0]
    r.ops, r.nops, r.segments, r.nseg = ops.ctypes.data, len(ops), segs.ctypes.data, len(segs)
    r.passes, r.npasses = arr.ctypes.data, int((segs[:, 0] == 1).sum())
    r.d_T0, r.d_T1, r.peak_bits = 1 << 20, 1 << 21, plan.peak_bits      # never dereferenced here
    nb = ctypes.c_uint64(0)
    assert L.tdp_run_programs_workspace(runs, 1, ctypes.byref(nb)) == 0 and nb.value > 0
    arr[0]["runs"][0, 0] = 0 | 1 << 8 | 38 << 16
    arr[0]["nruns"][0] = max(1, int(arr[0]["nruns"][0]))
    # fake device pointers: the call must fail in host validation, before it touches them
    assert L.tdp_run_programs(runs, 1, 1 << 22, nb.value, None, 0, None) == 1
