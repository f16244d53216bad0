"""ctypes binding of libqbhip (include/qb.h).

The library is loaded from the package directory (built in-tree by
`make -C quantumbilliards.jl_amd/csrc`).  There is no fallback: if the shared
library is missing the import of any entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libqbhip.so"))

QB_OK, QB_E_ARG, QB_E_HIP, QB_E_NOMEM, QB_E_STATE, QB_E_NUMERIC = 0, -1, -2, -3, -4, -5
QB_CFIE_KRESS, QB_DLP_KRESS = 1, 2
QB_FP64, QB_MIXED, QB_MIXED_INT8RES = 0, 1, 2
QB_OUT_DEVICE = 1
QB_DERIV_REFERENCE_SIGN = 2
QB_ASM_CONJ_PAIRS = 4
QB_DERIV_MATRIX_SIGN = 8

EXPORTS = (
    "qb_ctx_create", "qb_ctx_destroy", "qb_last_error", "qb_ctx_device_bytes",
    "qb_geom_create", "qb_geom_destroy", "qb_geom_size", "qb_assemble",
    "qb_window_create", "qb_window_destroy", "qb_window_footprint", "qb_window_moments",
    "qb_bessel_eval", "qb_cheb_tune", "qb_ctx_set_profiling", "qb_ctx_kernel_stats",
    "qb_ctx_autotune", "qb_ctx_set_cache_limit", "qb_residuals", "qb_assemble_deriv",
    "qb_ctx_mem_info", "qb_ctx_set_graphs", "qb_window_moments_qr", "qb_window_basis",
    "qb_geom_set_orbits", "qb_geom_dim", "qb_host_alloc", "qb_host_free",
    "qb_geom_set_orbits_multi", "qb_geom_irreps", "qb_window_basis_irrep",
    "qb_window_create_multi", "qb_window_blocks", "qb_window_matrix_cols",
    "qb_ctx_kernel_issued", "qb_assemble_noalloc",
)
KERNEL_CLASSES = ("tables", "assemble", "gemm_c64", "gemm_c128", "panel", "trsm", "other",
                  "ozaki", "gemm_c64_bf16")


class QBError(RuntimeError):
    """A nonzero status from libqbhip, with the context's message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libqbhip error {code}: {msg}")
        self.code = code


class QBArgError(QBError, AssertionError):
    """QB_E_ARG: an argument the library rejects (sizes, ranges, operator,
    geometry kind).  Also an AssertionError, the exception the reference
    raises for the same checks (cheb-CFIE:1433, cheb-DLP:1117)."""


class ChebParams(C.Structure):
    _fields_ = [("n_panels_h", C.c_int32), ("M_h", C.c_int32),
                ("n_panels_j", C.c_int32), ("M_j", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("N", C.c_int64), ("nq", C.c_int32), ("r", C.c_int32),
                ("precision", C.c_int32), ("refine_iters_max", C.c_int32),
                ("nodes_fp64_fallback", C.c_int32), ("refine_resid_max", C.c_double),
                ("t_assemble_ms", C.c_double), ("t_factor_ms", C.c_double),
                ("t_solve_ms", C.c_double), ("rmin", C.c_double), ("rmax", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libqbhip once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libqbhip.so not built at {LIB_PATH}: run __graft_entry__.build() "
                      "or make -C quantumbilliards.jl_amd/csrc")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    PP = C.POINTER(C.c_void_p)
    sig = {
        "qb_ctx_create": (C.c_int, [C.c_int, PP]),
        "qb_ctx_destroy": (None, [P]),
        "qb_last_error": (C.c_char_p, [P]),
        "qb_ctx_device_bytes": (I64, [P]),
        "qb_geom_create": (C.c_int, [P, I32, P, P, P, P, P, P, P, P, P, PP]),
        "qb_geom_destroy": (None, [P]),
        "qb_geom_size": (I64, [P]),
        "qb_geom_dim": (I64, [P]),
        "qb_host_alloc": (C.c_int, [P, I64, PP]),
        "qb_host_free": (None, [P]),
        "qb_geom_set_orbits": (C.c_int, [P, P, I64, I32, P, P]),
        "qb_geom_set_orbits_multi": (C.c_int, [P, P, I64, I32, I32, P, P]),
        "qb_geom_irreps": (I32, [P]),
        "qb_window_basis_irrep": (C.c_int, [P, I32, P, I32, P]),
        "qb_assemble": (C.c_int, [P, P, I32, P, I64, C.POINTER(ChebParams), P, I32, P]),
        "qb_assemble_noalloc": (C.c_int, [P, P, I32, P, I64, C.POINTER(ChebParams), P, I32, P]),
        "qb_assemble_deriv": (C.c_int, [P, P, I32, P, I64, C.POINTER(ChebParams), P, P, P, I32,
                                        P]),
        "qb_window_create": (C.c_int, [P, P, I32, D, D, D, I32, C.POINTER(ChebParams), I32, PP,
                                       C.POINTER(Stats)]),
        "qb_window_create_multi": (C.c_int, [P, P, I32, I32, P, P, I32, C.POINTER(ChebParams),
                                             I32, PP, C.POINTER(Stats)]),
        "qb_window_blocks": (I32, [P]),
        "qb_window_matrix_cols": (C.c_int, [P, I32, I64, I64, P]),
        "qb_window_destroy": (None, [P]),
        "qb_window_footprint": (I64, [I64, I32, I32, I32]),
        "qb_window_moments": (C.c_int, [P, I32, P, P, P, D, C.POINTER(Stats)]),
        "qb_window_moments_qr": (C.c_int, [P, I32, P, P, P, D, C.POINTER(Stats)]),
        "qb_window_basis": (C.c_int, [P, P, I32, P]),
        "qb_bessel_eval": (C.c_int, [P, P, I64, P, P, P, P]),
        "qb_cheb_tune": (C.c_int, [P, P, I32, P, I64, D, I32, I32, C.POINTER(ChebParams), P]),
        "qb_ctx_set_profiling": (C.c_int, [P, I32]),
        "qb_ctx_autotune": (C.c_int, [P, I64, I32, I32, I32, C.c_char_p, I64]),
        "qb_ctx_set_cache_limit": (C.c_int, [P, I64]),
        "qb_ctx_mem_info": (C.c_int, [P, C.POINTER(I64), C.POINTER(I64)]),
        "qb_ctx_set_graphs": (C.c_int, [P, I32]),
        "qb_residuals": (C.c_int, [P, P, I32, P, I64, C.POINTER(ChebParams), P, I32, P, P, P, P,
                                   P]),
        "qb_ctx_kernel_stats": (C.c_int, [P, I32, C.POINTER(D), C.POINTER(I64), C.POINTER(D),
                                          C.POINTER(D)]),
        "qb_ctx_kernel_issued": (C.c_int, [P, I32, C.POINTER(D)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(ctx_ptr, rc: int):
    if rc != QB_OK:
        msg = lib().qb_last_error(ctx_ptr)
        cls = QBArgError if rc == QB_E_ARG else QBError
        raise cls(rc, msg.decode() if msg else "")
