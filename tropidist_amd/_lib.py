"""ctypes binding of libtdp.so (include/tdp.h) and the device-memory plumbing shared by the host modules.

The library must be built in-tree (python -m tropidist_amd.build); loading fails loudly when it is missing or when
no HIP device is visible -- this package has no CPU execution path."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtdp.so")

TDP_OK, TDP_EINVAL, TDP_EHIP, TDP_ENOMEM, TDP_EDEADLINE, TDP_ECAPACITY = range(6)

_LIB = None
vp, u64, i64, ci, u8 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_uint8

# name -> (restype, argtypes); every entry point returns an int status
SIGNATURES = {
    "tdp_last_error": (ctypes.c_char_p, []),
    "tdp_version": (ctypes.c_char_p, []),
    "tdp_num_cus": (ci, [ctypes.POINTER(ci)]),
    "tdp_update": (ci, [vp, u64, u64, vp]),
    "tdp_update_w": (ci, [vp, u64, u64, ci, vp]),
    "tdp_expand_update": (ci, [vp, u64, u64, ci, vp]),
    "tdp_select": (ci, [vp, vp, u64, ci, ci, ctypes.POINTER(u64), vp]),
    "tdp_fill": (ci, [vp, u64, u8, vp]),
    "tdp_hist": (ci, [vp, u64, vp, vp]),
    "tdp_run_plan": (ci, [vp, i64, ci, vp, vp, ctypes.POINTER(ci), vp]),
    "tdp_plan_small_bits": (ci, []),
    "tdp_plan_small_bytes": (ci, [i64, ctypes.POINTER(u64)]),
    "tdp_run_plan_small": (ci, [vp, i64, ci, vp, u64, ctypes.POINTER(ci), vp]),
    "tdp_pass_args_bytes": (ci, []),
    "tdp_run_pass": (ci, [vp, vp, vp, vp]),
    "tdp_witness_step_bytes": (ci, []),
    "tdp_run_record_bytes": (ci, []),
    "tdp_read_back": (ci, [vp, vp, u64, vp]),
    "tdp_rank_profile": (ci, [vp, ci, ci, vp, ci, vp]),
    "tdp_run_programs_workspace": (ci, [vp, ci, ctypes.POINTER(u64)]),
    "tdp_run_programs": (ci, [vp, ci, vp, u64, vp, u64, vp]),
    "tdp_programs_graph": (vp, [vp, ci, vp, u64, vp]),
    "tdp_graph_launch": (ci, [vp, vp]),
    "tdp_graph_values": (ci, [vp, vp, ci, vp]),
    "tdp_graph_nodes": (ci, [vp, ctypes.POINTER(u64)]),
    "tdp_graph_free": (None, [vp]),
    "tdp_run_programs_timed": (ci, [vp, ci, vp, u64, vp, u64, vp, vp, vp, i64, ctypes.POINTER(i64)]),
    "tdp_fused_compile": (ci, [vp, i64, ci, ci, ci, vp, i64, ctypes.POINTER(i64), vp, i64, ctypes.POINTER(i64), vp,
                               i64, ctypes.POINTER(i64)]),
    "tdp_program_witness_bytes": (ci, [vp, i64, ctypes.POINTER(u64)]),
    "tdp_run_program_witness": (ci, [vp, i64, vp, i64, vp, i64, vp, u64, ctypes.POINTER(ci), vp,
                                     ctypes.POINTER(i64), vp]),
    "tdp_run_program": (ci, [vp, i64, vp, i64, vp, i64, ci, vp, vp, ctypes.POINTER(ci), vp]),
    "tdp_witness_workspace_bytes": (ci, [i64, ci, ctypes.POINTER(u64)]),
    "tdp_run_plan_witness": (ci, [vp, i64, ci, vp, u64, ctypes.POINTER(ci), vp, ctypes.POINTER(i64), vp]),
    "tdp_tree_join": (ci, [ci, vp, vp, vp, vp, vp, ctypes.c_uint, ctypes.c_uint, u64, u64, u64, vp]),
    "tdp_tree_tile_plan": (ci, [vp, vp, ctypes.c_uint, ctypes.c_uint, ci, u64, ci, vp, vp]),
    "tdp_tree_tile_args_bytes": (ci, []),
    "tdp_tree_join_u8": (ci, [vp, vp, vp, vp, vp, ctypes.c_uint, ctypes.c_uint, u64, u64, u64, vp]),
    "tdp_tree_join_u16": (ci, [vp, vp, vp, vp, vp, ctypes.c_uint, ctypes.c_uint, u64, u64, u64, vp]),
    "tdp_tree_join_u32": (ci, [vp, vp, vp, vp, vp, ctypes.c_uint, ctypes.c_uint, u64, u64, u64, vp]),
    "tdp_tree_join_u64": (ci, [vp, vp, vp, vp, vp, ctypes.c_uint, ctypes.c_uint, u64, u64, u64, vp]),
    "tdp_tree_move": (ci, [vp, vp, ctypes.c_size_t, vp]),
    "tdp_tree_leaf": (ci, [ci, vp, u64, u64, u64, u64, vp]),
    "tdp_tree_fill": (ci, [ci, vp, u64, u64, vp]),
    "tdp_tree_compile": (vp, [vp, ci, ci, vp, vp, vp, ci, ctypes.c_double, ctypes.POINTER(ci)]),
    "tdp_tree_free": (None, [vp]),
    "tdp_tree_info": (ci, [vp, vp]),
    "tdp_tree_nodes": (ci, [vp, vp, vp, vp, vp]),
    "tdp_tree_maps": (ci, [vp, ci, vp, vp]),
    "tdp_tree_interface": (ci, [vp, ci, vp]),
    "tdp_tree_solve": (ci, [vp, vp, vp, ctypes.POINTER(ci)]),
    "tdp_tree_run": (ci, [vp, vp, ci, vp, u64, u64, ctypes.c_double, ctypes.POINTER(u64), vp, vp]),
    "tdp_sp_index": (ci, [vp, u64, vp, u64, vp]),
    "tdp_sp_step": (ci, [vp, u64, vp, u64, vp, vp, vp, ci, ci, ci, vp, vp, i64, vp, ci, ci, vp, u64, vp, u64, vp,
                         vp]),
    "tdp_sp_lookup": (ci, [vp, vp, u64, vp, u64, vp, vp]),
    "tdp_sp_truncate": (ci, [vp, u64, ci, vp, vp, vp]),
    "tdp_sp_join": (ci, [vp, u64, vp, vp, u64, vp, vp, vp]),
    "tdp_sp_hist": (ci, [vp, u64, vp, vp]),
    "tdp_dn_from_sparse": (ci, [vp, u64, vp, ci, vp, vp]),
    "tdp_dn_to_sparse": (ci, [vp, ci, vp, vp, u64, vp, vp]),
    "tdp_dn_step": (ci, [vp, ci, vp, ci, vp, ci, vp, vp, ci, u64, ci, ci, ci, vp, vp, i64, u64, ci, vp, vp]),
    "tdp_dn_join": (ci, [vp, u64, vp, ci, vp, vp, vp, vp]),
    "tdp_tab_create": (vp, [ci, ci]),
    "tdp_tab_free": (None, [vp]),
    "tdp_tab_count": (i64, [vp]),
    "tdp_tab_build": (ci, [vp, ci, vp, vp, vp, ci]),
    "tdp_tab_build_noalloc": (ci, [vp, ci, vp, vp, vp, ci]),
    "tdp_tab_lookup_batch": (ci, [vp, vp, vp, vp, vp, u64, vp, vp]),
    "tdp_gm_set_cover": (None, [ci, ci]),
    "tdp_gm_set_kernel": (None, [ci, vp, vp, vp, vp]),
    "tdp_gm_set_kernel_mode": (None, [ci]),
    "tdp_gm_set_full_labels": (None, [vp, ci]),
    "tdp_gm_set_pattern": (None, [vp, vp, ci, ci]),
    "tdp_gm_num_threads": (ci, []),
    "tdp_gm_set_chunk": (None, [ci]),
    "tdp_gm_last_launches": (ci, []),
    "tdp_gm_release": (None, []),
    "tdp_gm_search": (ci, [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, vp, vp, vp, ci, vp, ci, vp, vp, vp, ci, ci, ci,
                           ctypes.c_double, vp, vp, vp, ci, ci, vp, vp, vp]),
    "tdp_batch_run": (ci, [vp, vp, ci, vp, ci, vp]),
    "tdp_batch_item_bytes": (ci, []),
    "tdp_plan_build": (ci, [vp, ci, ci, vp, vp, ci, vp, vp, i64, vp, vp, vp, vp, vp]),
    "tdp_greedy_order": (ci, [vp, ci, ci, ci, vp]),
    "tdp_sp_trellis": (ci, [vp, ci, ci, vp, vp, ci, ci, vp, vp, vp, vp]),
    "tdp_sq_run": (ci, [vp, vp, vp, vp, vp, vp, vp]),
    "tdp_sp_bounds": (ci, [vp, ci, ci, vp, vp, ci, ci, vp, vp, vp, vp]),
    "tdp_spb_create": (vp, [u64, ci, vp]),
    "tdp_spb_free": (ci, [vp]),
    "tdp_spb_add": (ci, [vp, vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, u64, ctypes.POINTER(i64)]),
    "tdp_spb_step": (ci, [vp, ctypes.POINTER(ctypes.c_int32)]),
    "tdp_spb_stats": (ci, [vp, vp]),
    "tdp_spb_poll": (ci, [vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "tdp_spb_result": (ci, [vp, i64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), vp]),
    "tdp_spb_witness": (ci, [vp, i64, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "tdp_spb_release": (ci, [vp, i64]),
    "tdp_spb_cancel": (ci, [vp, i64]),
}


class TdpError(RuntimeError):
    """A libtdp call failed; .code is its status."""

    def __init__(self, message, code):
        super().__init__(message)
        self.code = code


def lib():
    """Load libtdp.so.  torch is imported first: its HIP runtime (SONAME libamdhip64.so.7) is then the one the
    library binds to, so the process holds a single HIP runtime and device memory is shared with torch tensors."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtdp.so is not built: run `python -m tropidist_amd.build` (no CPU fallback)")
        import torch  # noqa: F401  (loads the HIP runtime libtdp.so links against)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != TDP_OK:
        msg = lib().tdp_last_error().decode(errors="replace")
        if rc == TDP_EDEADLINE:
            raise TimeoutError("%s: %s" % (what, msg))
        if rc == TDP_ENOMEM:
            raise MemoryError("%s: %s" % (what, msg))
        raise TdpError("%s failed (status %d): %s" % (what, rc, msg), rc)


# ------------------------------------------------------------------------------------------ device plumbing
_torch = None


def torch():
    """torch with a visible HIP device (device memory and streams only)."""
    global _torch
    if _torch is None:
        import torch as t
        if not t.cuda.is_available():
            raise RuntimeError("tropidist_amd needs a HIP device (torch.cuda.is_available() is False)")
        _torch = t
    return _torch


def retry_after_release(fn):
    """fn(), and once more after torch's cached device blocks are handed back when it raises MemoryError: the
    library's own allocations (lookup tables, the search's node buffers, the sparse engine's arena) come from the
    driver, which does not see memory that torch holds cached for later tensors."""
    try:
        return fn()
    except MemoryError:
        t = torch()
        t.cuda.synchronize()
        t.cuda.empty_cache()
        return fn()


def device_empty(nbytes, dtype=None):
    t = torch()
    return t.empty(max(1, int(nbytes)), dtype=dtype or t.uint8, device="cuda")


def stream_handle(stream=None):
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def dptr(tensor):
    return ctypes.c_void_p(tensor.data_ptr())


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p)
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)


class TorchDeviceAllocator:
    """Device buffers for a native call, taken from torch's caching allocator on the stream the call runs on, as a
    pair of C callbacks (alloc(bytes, ctx) -> pointer or NULL, free(pointer, ctx)).  Memory a call gives back stays in
    torch's cache for the next one instead of going back to the driver.  Use as a context manager around the call:
    on exit every buffer the call did not free is released.  The callbacks run on the calling thread."""

    def __init__(self, stream=None):
        t = torch()
        self._stream = stream if stream is not None else t.cuda.current_stream()
        self._live = {}
        self.alloc_fn = ALLOC_FN(self._alloc)
        self.free_fn = FREE_FN(self._free)

    def _alloc(self, nbytes, ctx):
        t = torch()
        try:
            with t.cuda.stream(self._stream):
                buf = t.empty(max(1, int(nbytes)), dtype=t.uint8, device="cuda")
        except Exception:                       # out of memory: the native call reports it
            return None
        p = buf.data_ptr()
        self._live[p] = buf
        return p

    def _free(self, p, ctx):
        self._live.pop(p, None)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self._live.clear()
        return False
