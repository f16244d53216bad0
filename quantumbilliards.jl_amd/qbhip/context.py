"""Device context (qb_ctx): one device, one stream."""
from __future__ import annotations

import ctypes as C
import threading
import weakref

import numpy as np

from . import _lib


class Context:
    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = C.c_void_p()
        rc = L.qb_ctx_create(int(device), C.byref(h))
        if rc != _lib.QB_OK:
            raise _lib.QBError(rc, f"qb_ctx_create(device={device}) failed (no usable GPU?)")
        self.ptr = h
        self.device = device

    def device_bytes(self) -> int:
        return int(_lib.lib().qb_ctx_device_bytes(self.ptr))

    def set_profiling(self, enable: bool = True):
        """Record HIP events around every launch (and clear the counters)."""
        _lib.check(self.ptr, _lib.lib().qb_ctx_set_profiling(self.ptr, 1 if enable else 0))

    def kernel_stats(self) -> dict:
        """Per kernel class: device ms, launches, algorithmic flops and bytes,
        and the operations issued to the matrix cores (qb_ctx_kernel_issued)."""
        out = {}
        for k, name in enumerate(_lib.KERNEL_CLASSES):
            ms, n, fl, by = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
            _lib.check(self.ptr, _lib.lib().qb_ctx_kernel_stats(self.ptr, k, C.byref(ms), C.byref(n),
                                                                C.byref(fl), C.byref(by)))
            iss = C.c_double()
            _lib.check(self.ptr, _lib.lib().qb_ctx_kernel_issued(self.ptr, k, C.byref(iss)))
            out[name] = {"ms": ms.value, "launches": n.value, "flops": fl.value, "bytes": by.value,
                         "issued": iss.value}
        return out

    def set_graphs(self, on: bool):
        """Replay the batched LU of windows through cached hipGraphs."""
        _lib.check(self.ptr, _lib.lib().qb_ctx_set_graphs(self.ptr, 1 if on else 0))

    def mem_info(self):
        """(free, total) device bytes of this context's GPU."""
        f, t = C.c_int64(), C.c_int64()
        _lib.check(self.ptr, _lib.lib().qb_ctx_mem_info(self.ptr, C.byref(f), C.byref(t)))
        return f.value, t.value

    def set_cache_limit(self, nbytes: int):
        """Bytes of freed device buffers kept for reuse (0 disables caching)."""
        _lib.check(self.ptr, _lib.lib().qb_ctx_set_cache_limit(self.ptr, int(nbytes)))

    def autotune(self, N: int, nq: int, r: int, precision: str = "mixed") -> str:
        """Pick GEMM tile variants for the shapes of an (N, nq, r) window
        (qb_ctx_autotune); returns the report."""
        buf = C.create_string_buffer(8192)
        prec = {"fp64": _lib.QB_FP64, "mixed": _lib.QB_MIXED,
                "mixed_int8res": _lib.QB_MIXED_INT8RES}[precision]
        _lib.check(self.ptr, _lib.lib().qb_ctx_autotune(self.ptr, int(N), int(nq), int(r), prec,
                                                        buf, len(buf)))
        return buf.value.decode()

    def close(self):
        if getattr(self, "ptr", None):
            _lib.lib().qb_ctx_destroy(self.ptr)
            self.ptr = None


_default = {}
_lock = threading.Lock()


def default_context(device: int = 0) -> Context:
    with _lock:
        c = _default.get(device)
        if c is None:
            c = Context(device)
            _default[device] = c
        return c


def pinned_array(shape, dtype=np.complex128, order: str = "F", ctx: "Context" = None) -> np.ndarray:
    """An uninitialised NumPy array in page-locked host memory (qb_host_alloc):
    the library moves it to and from the device by DMA.  The memory is
    released when the last view of the array is gone."""
    ctx = ctx or default_context()
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    p = C.c_void_p()
    _lib.check(ctx.ptr, _lib.lib().qb_host_alloc(ctx.ptr, max(n, 1), C.byref(p)))
    buf = (C.c_char * max(n, 1)).from_address(p.value)
    weakref.finalize(buf, _lib.lib().qb_host_free, C.c_void_p(p.value))
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape, order=order)
