"""ctypes binding of libsfmi355x.so (include/sfmi355x.h).

The library is built in-tree (structurefunctions.jl_amd/libsfmi355x.so).  There is no fallback:
if it cannot be loaded, every entry point raises BackendError.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import sys

import numpy as np

from .errors import ArgumentError, BackendError, DimensionMismatch

LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libsfmi355x.so")

SF_OK, SF_E_ARG, SF_E_DIM, SF_E_COUNT, SF_E_DEVICE, SF_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5
PREC_FP64, PREC_MIXED, PREC_MIXED_FIXED = 0, 1, 2
SEM_POINT, SEM_BATCH = 0, 1
OP_SINGLE_PASS = 6

EXPORTED = ("sf_create", "sf_destroy", "sf_last_error", "sf_hist", "sf_hist_batch",
            "sf_hist_device", "sf_hist_batch_device", "sf_last_stats", "sf_device_info", "sf_check",
            "sf_eval_math",
            "sf_set_tuning", "sf_plan_create", "sf_plan_execute", "sf_plan_sync", "sf_plan_destroy",
            "sf_autotune", "sf_hist_batch_device_noalloc", "sf_hist_multi_device", "sf_hist_multi",
            "sf_plan_create_multi", "sf_plan_execute_multi")
SF_MAX_FUSED = 8


class SfBinRule(ctypes.Structure):
    """sf_bin_rule: a LinearBinEdges axis (kind 1), optionally InfPadded."""
    _fields_ = [("kind", ctypes.c_int32), ("padded", ctypes.c_int32), ("n", ctypes.c_int64),
                ("first_edge", ctypes.c_double), ("last_edge", ctypes.c_double),
                ("inv_step", ctypes.c_double), ("step_val", ctypes.c_double)]


def make_rule(fields):
    """SfBinRule from the dict of sfamd.bins.rule_of (None stays None)."""
    if fields is None:
        return None
    return SfBinRule(1, 1 if fields["padded"] else 0, int(fields["n"]), float(fields["first_edge"]),
                     float(fields["last_edge"]), float(fields["inv_step"]), float(fields["step_val"]))


class SfProblem(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_int32), ("W", ctypes.c_int32), ("N", ctypes.c_int64),
        ("x", ctypes.c_void_p), ("u", ctypes.c_void_p),
        ("geometry", ctypes.c_int32), ("radius", ctypes.c_double),
        ("op", ctypes.c_int32), ("nl", ctypes.c_int32), ("nt", ctypes.c_int32),
        ("n_dist", ctypes.c_int32), ("dist_edges", ctypes.c_void_p),
        ("n_val", ctypes.c_int32), ("val_edges", ctypes.c_void_p),
        ("precision", ctypes.c_int32), ("count_bits", ctypes.c_int32),
        ("semantics", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dist_rule", ctypes.c_void_p), ("val_rule", ctypes.c_void_p),
    ]


class SfOutput(ctypes.Structure):
    """sf_output: one operator of a fused call (n_val > 0: the joint histogram's value bins)."""
    _fields_ = [("op", ctypes.c_int32), ("nl", ctypes.c_int32), ("nt", ctypes.c_int32),
                ("n_val", ctypes.c_int32), ("val_edges", ctypes.c_void_p), ("val_rule", ctypes.c_void_p)]


_lib = None


def _share_torch_hip_runtime():
    """Make this library bind to the HIP runtime PyTorch-ROCm uses, when torch is installed.

    torch ships its own libamdhip64 / libhsa-runtime64 (same sonames as /opt/rocm's) and links them
    by file name.  If this library were loaded first, the process would hold /opt/rocm's runtime and
    torch would then load its own copy, which cannot open the GPU a second time.  Loading torch's
    copy first makes both resolve to one runtime whatever the import order.  Set
    SFMI355X_OWN_HIP_RUNTIME=1 to skip this (then import torch before this library).
    """
    if os.environ.get("SFMI355X_OWN_HIP_RUNTIME") or "torch" in sys.modules:
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    hip = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(hip):
        ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)


def lib():
    """Load the library once; raise BackendError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C csrc")
    _share_torch_hip_runtime()
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    L.sf_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.sf_create.restype = ctypes.c_int
    L.sf_destroy.argtypes = [vp]
    L.sf_destroy.restype = None
    L.sf_last_error.argtypes = [vp]
    L.sf_last_error.restype = ctypes.c_char_p
    L.sf_hist.argtypes = [vp, ctypes.POINTER(SfProblem), dp, dp]
    L.sf_hist.restype = ctypes.c_int
    L.sf_hist_batch.argtypes = [vp, i32, ctypes.POINTER(SfProblem), dp, dp]
    L.sf_hist_batch.restype = ctypes.c_int
    L.sf_hist_device.argtypes = [vp, ctypes.POINTER(SfProblem), dp, dp, vp]
    L.sf_hist_device.restype = ctypes.c_int
    L.sf_hist_batch_device.argtypes = [vp, i32, ctypes.POINTER(SfProblem), dp, dp, vp]
    L.sf_hist_batch_device.restype = ctypes.c_int
    L.sf_last_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.sf_last_stats.restype = ctypes.c_int
    L.sf_device_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.sf_device_info.restype = ctypes.c_int
    L.sf_check.argtypes = [i32, ctypes.POINTER(SfProblem), ctypes.c_char_p, i32]
    L.sf_check.restype = ctypes.c_int
    L.sf_eval_math.argtypes = [vp, i32, i64, dp, dp, dp]
    L.sf_eval_math.restype = ctypes.c_int
    L.sf_set_tuning.argtypes = [vp, i32, i32, i32, i32, i32, i32]
    L.sf_set_tuning.restype = ctypes.c_int
    L.sf_plan_create.argtypes = [vp, i32, ctypes.POINTER(SfProblem), ctypes.POINTER(vp)]
    L.sf_plan_create.restype = ctypes.c_int
    L.sf_plan_execute.argtypes = [vp, dp, dp, vp, i32]
    L.sf_plan_execute.restype = ctypes.c_int
    L.sf_plan_sync.argtypes = [vp, ctypes.POINTER(i32)]
    L.sf_plan_sync.restype = ctypes.c_int
    L.sf_plan_destroy.argtypes = [vp]
    L.sf_plan_destroy.restype = None
    L.sf_autotune.argtypes = [vp, i32, ctypes.POINTER(SfProblem), ctypes.POINTER(i32)]
    L.sf_autotune.restype = ctypes.c_int
    L.sf_hist_batch_device_noalloc.argtypes = [vp, i32, ctypes.POINTER(SfProblem), dp, dp, vp]
    L.sf_hist_batch_device_noalloc.restype = ctypes.c_int
    L.sf_hist_multi_device.argtypes = [vp, i32, ctypes.POINTER(SfProblem), i32, ctypes.POINTER(SfOutput),
                                       ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
    L.sf_hist_multi_device.restype = ctypes.c_int
    L.sf_hist_multi.argtypes = [vp, i32, ctypes.POINTER(SfProblem), i32, ctypes.POINTER(SfOutput),
                                ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.sf_hist_multi.restype = ctypes.c_int
    L.sf_plan_create_multi.argtypes = [vp, i32, ctypes.POINTER(SfProblem), i32, ctypes.POINTER(SfOutput),
                                       ctypes.POINTER(vp)]
    L.sf_plan_create_multi.restype = ctypes.c_int
    L.sf_plan_execute_multi.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp, i32]
    L.sf_plan_execute_multi.restype = ctypes.c_int
    _lib = L
    return L


def raise_for(code: int, msg: str):
    if code == SF_OK:
        return
    if code == SF_E_DIM:
        raise DimensionMismatch(msg)
    if code in (SF_E_ARG, SF_E_COUNT, SF_E_UNSUPPORTED):
        raise ArgumentError(msg)
    raise BackendError(msg or f"libsfmi355x error {code}")


class Context:
    """One device context (sf_create); not thread-safe, one call at a time."""

    def __init__(self, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.sf_create(int(device), ctypes.byref(h))
        if rc != SF_OK:
            raise BackendError(f"sf_create(device={device}) failed with code {rc} (no usable GPU?)")
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self) -> str:
        m = lib().sf_last_error(self._h)
        return m.decode() if m else ""

    def check(self, rc: int):
        if rc != SF_OK:
            raise_for(rc, self.error())

    def device_info(self):
        """Compute units, peak shader clock (kHz), LDS bytes per workgroup, lanes per wavefront."""
        out = (ctypes.c_int64 * 4)()
        self.check(lib().sf_device_info(self._h, out))
        return {"cus": out[0], "clock_khz": out[1], "lds_bytes": out[2], "wave": out[3]}

    def stats(self):
        out = (ctypes.c_int64 * 8)()
        lib().sf_last_stats(self._h, out)
        return {"pairs": out[0], "rebinned": out[1], "items": out[2], "blocks": out[3],
                "hist_ns": out[4], "recheck_ns": out[5], "call_ns": out[6], "block": out[7]}

    def set_tuning(self, block=0, j_chunk=0, blocks_per_cu=0, variant=0, copies=0, layout=0):
        self.check(lib().sf_set_tuning(self._h, int(block), int(j_chunk), int(blocks_per_cu), int(variant),
                                       int(copies), int(layout)))

    def autotune(self, problems):
        """sf_autotune on device-pointer problems; returns (block, j_chunk, blocks_per_cu, variant,
        copies, layout)."""
        arr = (SfProblem * len(problems))(*problems)
        out = (ctypes.c_int32 * 6)()
        self.check(lib().sf_autotune(self._h, len(problems), arr, out))
        return tuple(out)

    def eval_math(self, fn: int, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = len(a)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros(2 * n if fn <= 1 else n)
        rc = lib().sf_eval_math(self._h, fn, n, a.ctypes.data, None if bb is None else bb.ctypes.data,
                                out.ctypes.data)
        self.check(rc)
        return out.reshape(n, 2) if fn <= 1 else out

    def hist(self, problems, sums, counts):
        """Host-pointer batch call; sums/counts are contiguous arrays (n_slices * ncell)."""
        arr = (SfProblem * len(problems))(*problems)
        rc = lib().sf_hist_batch(self._h, len(problems), arr, sums.ctypes.data, counts.ctypes.data)
        self.check(rc)

    def hist_multi(self, problems, outputs, sums_list, counts_list):
        """sf_hist_multi: host-pointer fused call; sums_list[k] / counts_list[k] are contiguous
        float64 / uint64 arrays of n_slices * ncell_k, accumulated into."""
        arr = (SfProblem * len(problems))(*problems)
        outs = (SfOutput * len(outputs))(*outputs)
        ps = (ctypes.c_void_p * len(outputs))(*[a.ctypes.data for a in sums_list])
        pc = (ctypes.c_void_p * len(outputs))(*[a.ctypes.data for a in counts_list])
        self.check(lib().sf_hist_multi(self._h, len(problems), arr, len(outputs), outs, ps, pc))

    def hist_multi_device(self, problems, outputs, d_sums_ptrs, d_counts_ptrs, stream_ptr: int = 0):
        """sf_hist_multi_device: fused call on device pointers (one sums / counts address per output)."""
        arr = (SfProblem * len(problems))(*problems)
        outs = (SfOutput * len(outputs))(*outputs)
        ps = (ctypes.c_void_p * len(outputs))(*[int(a) for a in d_sums_ptrs])
        pc = (ctypes.c_void_p * len(outputs))(*[int(a) for a in d_counts_ptrs])
        self.check(lib().sf_hist_multi_device(self._h, len(problems), arr, len(outputs), outs, ps, pc,
                                              ctypes.c_void_p(stream_ptr) if stream_ptr else None))

    def hist_device_noalloc(self, problems, d_sums_ptr: int, d_counts_ptr: int, stream_ptr: int = 0):
        """sf_hist_batch_device_noalloc: as hist_device, with no device allocation."""
        arr = (SfProblem * len(problems))(*problems)
        rc = lib().sf_hist_batch_device_noalloc(self._h, len(problems), arr, ctypes.c_void_p(d_sums_ptr),
                                                ctypes.c_void_p(d_counts_ptr),
                                                ctypes.c_void_p(stream_ptr) if stream_ptr else None)
        self.check(rc)

    def hist_device(self, problems, d_sums_ptr: int, d_counts_ptr: int, stream_ptr: int = 0):
        """Device-pointer batch call (x/u in each problem and the outputs are device addresses)."""
        arr = (SfProblem * len(problems))(*problems)
        rc = lib().sf_hist_batch_device(self._h, len(problems), arr, ctypes.c_void_p(d_sums_ptr),
                                        ctypes.c_void_p(d_counts_ptr),
                                        ctypes.c_void_p(stream_ptr) if stream_ptr else None)
        self.check(rc)


class Plan:
    """sf_plan: a batch built once and executed many times (optionally as a hipGraph)."""

    def __init__(self, ctx: Context, problems, outputs=None):
        """outputs: None for a one-operator plan, else a list of SfOutput (a fused plan)."""
        self.ctx = ctx
        self._problems = (SfProblem * len(problems))(*problems)
        self.n_out = 0 if outputs is None else len(outputs)
        h = ctypes.c_void_p()
        if outputs is None:
            ctx.check(lib().sf_plan_create(ctx.handle, len(problems), self._problems, ctypes.byref(h)))
        else:
            self._outs = (SfOutput * len(outputs))(*outputs)
            ctx.check(lib().sf_plan_create_multi(ctx.handle, len(problems), self._problems, len(outputs),
                                                 self._outs, ctypes.byref(h)))
        self._h = h

    def execute_multi(self, d_sums_ptrs, d_counts_ptrs, stream_ptr: int = 0, graph: bool = True):
        ps = (ctypes.c_void_p * self.n_out)(*[int(a) for a in d_sums_ptrs])
        pc = (ctypes.c_void_p * self.n_out)(*[int(a) for a in d_counts_ptrs])
        self.ctx.check(lib().sf_plan_execute_multi(self._h, ps, pc, ctypes.c_void_p(stream_ptr) if stream_ptr else None,
                                                   1 if graph else 0))

    def execute(self, d_sums_ptr: int, d_counts_ptr: int, stream_ptr: int = 0, graph: bool = True):
        self.ctx.check(lib().sf_plan_execute(self._h, ctypes.c_void_p(d_sums_ptr), ctypes.c_void_p(d_counts_ptr),
                                             ctypes.c_void_p(stream_ptr) if stream_ptr else None, 1 if graph else 0))

    def sync(self) -> bool:
        r = ctypes.c_int32(0)
        self.ctx.check(lib().sf_plan_sync(self._h, ctypes.byref(r)))
        return bool(r.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_problems(problems) -> tuple[int, str]:
    """sf_check: validation without a device."""
    arr = (SfProblem * len(problems))(*problems)
    buf = ctypes.create_string_buffer(512)
    rc = lib().sf_check(len(problems), arr, buf, 512)
    return rc, buf.value.decode()
