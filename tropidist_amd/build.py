"""Build libtdp.so (all HIP sources of csrc/, gfx950) in-tree with hipcc.

    python -m tropidist_amd.build

The library is written next to this file; the Python modules load it from there and fail loudly when it is
missing (there is no CPU fallback)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtdp.so")
ARCH = os.environ.get("TDP_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
    # code objects v5: loadable by the HIP runtime torch ships as well as by the system one
    "-mcode-object-version=5",
    # packed FP32 instructions stay off in every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-8000:])


def build(verbose=False, jobs=None):
    """Compile every csrc/*.hip and *.cpp (always recompiles), one object per source in parallel, and link them
    into libtdp.so.  Returns the library path."""
    import shutil
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    inc = "-I" + os.path.join(HERE, "..", "include")
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    tmpdir = tempfile.mkdtemp(prefix="tdp_build_")
    try:
        objs = [os.path.join(tmpdir, os.path.basename(src) + ".o") for src in sources()]
        jobs = jobs or min(8, os.cpu_count() or 4)
        with ThreadPoolExecutor(jobs) as pool:
            list(pool.map(lambda so: _run([hipcc] + compile_flags + [inc, "-c", so[0], "-o", so[1]], verbose),
                          zip(sources(), objs)))
        tmp = LIB + ".tmp%d" % os.getpid()
        try:
            _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp], verbose)
        except Exception:
            if os.path.exists(tmp):
                os.unlink(tmp)
            raise
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
