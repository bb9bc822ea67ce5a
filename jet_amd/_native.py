"""ctypes binding of libjet_hip.so (the C ABI declared in include/jet_hip.h).

The library is built in-tree by ``python -m jet_amd.build``.  There is no
fallback: if the library is missing or no GPU is present, the calls that need
it raise ``JetHipError``.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libjet_hip.so")

C64 = 0
C128 = 1
ROW_MAJOR = 0
COL_MAJOR = 1

PREC_DEFAULT = 0
PREC_FP32 = 1
PREC_BF16X3 = 2
PREC_FP64 = 3
PREC_BF16X3_4M = 4
PREC_F16X3_4M = 5

GEMM_KERNEL_NAMES = {0: "valu", 1: "valu_reduce", 2: "fp32_mfma", 3: "fp64_mfma", 4: "bf16x3_mfma",
                     5: "bf16x3_4m_mfma", 6: "splitk", 7: "f16x3_4m_mfma"}

# Exported symbols (must match include/jet_hip.h).
EXPORTS = [
    "jet_hip_get_version", "jet_hip_last_error", "jet_hip_ctx_create", "jet_hip_ctx_destroy",
    "jet_hip_ctx_device", "jet_hip_ctx_stream", "jet_hip_malloc", "jet_hip_free",
    "jet_hip_memcpy", "jet_hip_memset", "jet_hip_stream_synchronize",
    "jet_hip_device_synchronize", "jet_hip_permute", "jet_hip_contraction_plan_create",
    "jet_hip_plan_workspace_size", "jet_hip_plan_info", "jet_hip_plan_destroy",
    "jet_hip_contract", "jet_hip_contract_noalloc", "jet_hip_add", "jet_hip_slice", "jet_hip_fill_random",
    "jet_hip_reduce_compensated", "jet_hip_convert", "jet_hip_event_create",
    "jet_hip_event_record", "jet_hip_event_elapsed_ms", "jet_hip_event_destroy",
    "jet_hip_mfma_ceiling", "jet_hip_profiler_enable", "jet_hip_profiler_reset",
    "jet_hip_profiler_read",
    "jet_hip_profiler_kernel_name", "jet_hip_network_create", "jet_hip_network_estimate", "jet_hip_network_info",
    "jet_hip_network_step_info", "jet_hip_network_execute", "jet_hip_network_capture",
    "jet_hip_network_launch", "jet_hip_network_destroy", "jet_hip_batch_capture",
    "jet_hip_batch_launch", "jet_hip_batch_destroy", "jet_hip_dag_create", "jet_hip_dag_estimate",
    "jet_hip_dag_value", "jet_hip_gemm_choice", "jet_hip_tuning_set", "jet_hip_tuning_get",
    "jet_hip_autotune_gemm",
]


class JetHipError(RuntimeError):
    """Raised for any failing libjet_hip call; the message is the library's
    Jet::Exception-style text."""


_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _declare(L):
    L.jet_hip_last_error.restype = ctypes.c_char_p
    L.jet_hip_last_error.argtypes = []
    L.jet_hip_profiler_kernel_name.restype = ctypes.c_char_p
    L.jet_hip_profiler_kernel_name.argtypes = [ctypes.c_int]
    sig = {
        "jet_hip_get_version": [ctypes.POINTER(ctypes.c_int)] * 3,
        "jet_hip_ctx_create": [ctypes.c_int, ctypes.POINTER(_vp)],
        "jet_hip_ctx_destroy": [_vp],
        "jet_hip_ctx_device": [_vp, ctypes.POINTER(ctypes.c_int)],
        "jet_hip_ctx_stream": [_vp, ctypes.POINTER(_vp)],
        "jet_hip_malloc": [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)],
        "jet_hip_free": [_vp, _vp],
        "jet_hip_memcpy": [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp, ctypes.c_int],
        "jet_hip_memset": [_vp, _vp, ctypes.c_int, ctypes.c_size_t, _vp],
        "jet_hip_stream_synchronize": [_vp, _vp],
        "jet_hip_device_synchronize": [_vp],
        "jet_hip_permute": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i64p, _i32p,
                            _vp, _vp, _vp],
        "jet_hip_contraction_plan_create": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, _i32p, _i64p, ctypes.c_int, _i32p,
                                            _i64p, ctypes.c_int, _i32p, ctypes.POINTER(_vp)],
        "jet_hip_plan_workspace_size": [_vp, ctypes.POINTER(ctypes.c_size_t)],
        "jet_hip_plan_info": [_vp, ctypes.POINTER(ctypes.c_double), _i64p, _i64p, _i64p, _i64p,
                              ctypes.POINTER(ctypes.c_int)],
        "jet_hip_plan_destroy": [_vp],
        "jet_hip_contract": [_vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp],
        "jet_hip_contract_noalloc": [_vp, _vp, _vp, _vp, _vp],
        "jet_hip_gemm_choice": [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "jet_hip_tuning_set": [ctypes.c_int, _i32p],
        "jet_hip_tuning_get": [ctypes.c_int, _i32p, ctypes.POINTER(ctypes.c_int)],
        "jet_hip_autotune_gemm": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                  ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                  _i32p, _i32p, ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_int)],
        "jet_hip_add": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i64p, _i32p, _vp,
                        _vp, _vp, _vp],
        "jet_hip_slice": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int,
                          ctypes.c_int64, _vp, _vp, _vp],
        "jet_hip_fill_random": [_vp, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int,
                                _vp, _vp],
        "jet_hip_reduce_compensated": [_vp, ctypes.c_int, ctypes.c_size_t, _vp, _vp, _vp],
        "jet_hip_convert": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, _vp, _vp, _vp],
        "jet_hip_event_create": [_vp, ctypes.POINTER(_vp)],
        "jet_hip_event_record": [_vp, _vp, _vp],
        "jet_hip_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
        "jet_hip_event_destroy": [_vp],
        "jet_hip_network_create": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                   _i64p, ctypes.c_int, _i32p, ctypes.c_int, _i32p,
                                   ctypes.POINTER(_vp)],
        "jet_hip_network_estimate": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                     _i64p, ctypes.c_int, _i32p, ctypes.c_int, _i32p,
                                     ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_size_t),
                                     ctypes.POINTER(ctypes.c_size_t)],
        "jet_hip_network_info": [_vp, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                 _i64p, ctypes.POINTER(ctypes.c_int)],
        "jet_hip_network_step_info": [_vp, ctypes.c_int, _i64p, _i64p, _i64p, _i64p,
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "jet_hip_network_execute": [_vp, ctypes.POINTER(_vp), _vp, _vp],
        "jet_hip_network_capture": [_vp, ctypes.POINTER(_vp), _vp, _vp],
        "jet_hip_network_launch": [_vp, _vp],
        "jet_hip_network_destroy": [_vp],
        "jet_hip_batch_capture": [_vp, ctypes.c_int, ctypes.POINTER(_vp),
                                  ctypes.POINTER(ctypes.POINTER(_vp)), ctypes.POINTER(_vp), _vp,
                                  ctypes.POINTER(_vp)],
        "jet_hip_batch_launch": [_vp, _vp],
        "jet_hip_dag_create": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                               _i64p, ctypes.c_int, _i32p, ctypes.c_int, _i32p,
                               ctypes.POINTER(_vp)],
        "jet_hip_dag_estimate": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                 _i64p, ctypes.c_int, _i32p, ctypes.c_int, _i32p,
                                 ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_size_t),
                                 ctypes.POINTER(ctypes.c_size_t)],
        "jet_hip_dag_value": [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int),
                              _i32p, _i64p],
        "jet_hip_batch_destroy": [_vp],
        "jet_hip_mfma_ceiling": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double)],
        "jet_hip_profiler_enable": [ctypes.c_int],
        "jet_hip_profiler_reset": [],
        "jet_hip_profiler_read": [ctypes.c_int, _i64p, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int


def lib():
    """Load libjet_hip.so (raises JetHipError if it was not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise JetHipError(
                    f"{LIB_PATH} is missing: build it with `python -m jet_amd.build`")
            L = ctypes.CDLL(LIB_PATH)
            _declare(L)
            _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        msg = lib().jet_hip_last_error().decode(errors="replace")
        raise JetHipError(msg)


def i32(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (ctypes.c_int32 * max(1, len(vals)))(*vals)


def i64(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (ctypes.c_int64 * max(1, len(vals)))(*vals)


class Context:
    """A device context (owns one non-blocking stream)."""

    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().jet_hip_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        s = _vp()
        check(lib().jet_hip_ctx_stream(h, ctypes.byref(s)))
        self.stream = s

    def synchronize(self, stream=None) -> None:
        check(lib().jet_hip_stream_synchronize(self.handle, stream))

    def close(self) -> None:
        if self.handle:
            check(lib().jet_hip_ctx_destroy(self.handle))
            self.handle = None


_default_ctx: Optional[Context] = None


_ctx_lock = threading.Lock()


def default_context() -> Context:
    """The process-wide context of device JET_HIP_DEVICE (default 0), created
    once under a lock so concurrent first callers share it."""
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(int(os.environ.get("JET_HIP_DEVICE", "0")))
        return _default_ctx


class DeviceBuffer:
    """Owning device allocation."""

    def __init__(self, nbytes: int, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().jet_hip_malloc(self.ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, host, stream=None) -> None:
        import numpy as np
        host = np.ascontiguousarray(host)
        if self.ptr is None or host.nbytes > self.nbytes:
            raise JetHipError(f"Error in Jet: upload of {host.nbytes} bytes into a "
                              f"{self.nbytes}-byte device buffer.")
        check(lib().jet_hip_memcpy(self.ctx.handle, self.ptr, host.ctypes.data, host.nbytes, 0,
                                   stream, 0))

    def download(self, host, stream=None) -> None:
        if not host.flags["C_CONTIGUOUS"]:
            raise JetHipError("Error in Jet: download needs a C-contiguous host array.")
        if self.ptr is None or host.nbytes > self.nbytes:
            raise JetHipError(f"Error in Jet: download of {host.nbytes} bytes from a "
                              f"{self.nbytes}-byte device buffer.")
        check(lib().jet_hip_memcpy(self.ctx.handle, host.ctypes.data, self.ptr, host.nbytes, 1,
                                   stream, 0))

    def free(self) -> None:
        if self.ptr is not None and _lib is not None and self.ctx.handle:
            lib().jet_hip_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        h = _vp()
        check(lib().jet_hip_event_create(self.ctx.handle, ctypes.byref(h)))
        self.handle = h

    def record(self, stream=None):
        check(lib().jet_hip_event_record(self.ctx.handle, self.handle, stream))

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().jet_hip_event_elapsed_ms(self.handle, later.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if _lib is not None and self.handle:
                lib().jet_hip_event_destroy(self.handle)
        except Exception:
            pass


class Profiler:
    """Per-kernel-family device time and algorithmic work (HIP events)."""

    N = 32

    @staticmethod
    def enable(on: bool = True):
        check(lib().jet_hip_profiler_enable(1 if on else 0))

    @staticmethod
    def reset():
        check(lib().jet_hip_profiler_reset())

    @staticmethod
    def read() -> dict:
        n = Profiler.N
        launches = (ctypes.c_int64 * n)()
        ms = (ctypes.c_double * n)()
        flops = (ctypes.c_double * n)()
        nbytes = (ctypes.c_double * n)()
        count = ctypes.c_int()
        check(lib().jet_hip_profiler_read(n, launches, ms, flops, nbytes, ctypes.byref(count)))
        out = {}
        for k in range(count.value):
            if launches[k]:
                name = lib().jet_hip_profiler_kernel_name(k).decode()
                out[name] = dict(launches=int(launches[k]), ms=float(ms[k]),
                                 flops=float(flops[k]), bytes=float(nbytes[k]))
        return out
