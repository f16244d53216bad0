"""Beyn contour method on libqbhip.

Mirrors the reference entry points (src/solvers/acceleratedmethods/beyn.jl):
  construct_B_matrix(solver, pts, N, k0, R; nq=64, r=48, svd_tol=1e-14, rng,
                     use_chebyshev, n_panels_h, M_h, n_panels_j, M_j, info, multithreaded)
      -> (B, Uk)                                                    beyn.jl:257-322
  solve_vect(solver, basis, pts, k, dk; nq=32, r=48, ...) -> (lam, Uk, Y, k, dk, pts)
                                                                    beyn.jl:364-375
  solve(solver, basis, pts, k, dk; ...) -> lam                     beyn.jl:386-397
The contour nodes, assembly, LU of every node and the moment solves run on
the device (qb_window_*); the thin SVD with the absolute rank threshold,
B = U_k^H A1 W_k Sigma_k^{-1} and eig(B) run on the host in FP64 (LAPACK
through NumPy), as the reference runs them through LinearAlgebra.

Extensions (keyword arguments the reference does not have):
  precision="fp64"|"mixed"|"mixed_int8res": complex128 LU, or complex64 LU
      on MFMA with FP64 iterative refinement (refine_tol: normwise backward
      error at which a node stops; nodes that stagnate are refactored in
      FP64); the residual is a complex128 GEMM ("mixed") or is formed exactly
      on the int8 matrix cores ("mixed_int8res", 32 more bytes per matrix
      entry);
  grow_r: keep doubling r while the rank saturates (the reference retries
      once with 2r, beyn.jl:291-309, because of the unconditional break at
      :307); with grow_r=False the single retry is reproduced;
  V: explicit probe block instead of one drawn from rng.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .context import Context, default_context
from .solvers import Geometry, boundary_matrix_size, boundary_matrix_size_for_solver

_PREC = {"fp64": _lib.QB_FP64, "mixed": _lib.QB_MIXED, "mixed_int8res": _lib.QB_MIXED_INT8RES}


def probe_block(rng: np.random.Generator, N: int, r: int, pinned: bool = False) -> np.ndarray:
    """randn(rng, ComplexF64, N, r) (beyn.jl:193): real and imaginary parts
    N(0, 1/2), column-major; in page-locked host memory when pinned (its
    upload to the device is then a direct DMA)."""
    V = (rng.standard_normal((N, r)) + 1j * rng.standard_normal((N, r))) * np.sqrt(0.5)
    if not pinned:
        return np.asfortranarray(V)
    from .context import pinned_array
    P = pinned_array((N, r), np.complex128, "F")
    P[...] = V
    return P


class Window:
    """Contour nodes of one disk with T(z_j) assembled and factored on the
    device (qb_window).  `moments(V)` may be called repeatedly.

    k0 and R may also be sequences: the disks k0[c] + R[c] e^{i theta} on the
    same geometry then form one window whose matrices are assembled, factored
    and solved as one batch (qb_window_create_multi); moments come back as one
    block per (irrep, disk), block e * nwin + c."""

    def __init__(self, solver, pts, k0, R, nq: int, *, n_panels_h=15000, M_h=5,
                 n_panels_j=3000, M_j=5, precision="fp64", geom: Optional[Geometry] = None,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or (geom.ctx if geom is not None else default_context())
        self.geom = geom or Geometry(pts, self.ctx, solver)
        self.N = self.geom.dim  # reduced dimension with a solver symmetry
        self.nirr = getattr(self.geom, "nirr", 1)  # irreps factored together
        self.nq = nq
        multi = np.ndim(k0) > 0
        k0s = np.atleast_1d(np.asarray(k0, dtype=np.complex128))
        Rs = np.broadcast_to(np.asarray(R, dtype=np.float64), k0s.shape).copy()
        self.nwin = len(k0s)
        self.nblocks = self.nirr * self.nwin
        self.k0, self.R = (k0s, Rs) if multi else (complex(k0s[0]), float(Rs[0]))
        self.precision = precision
        cp = _lib.ChebParams(int(n_panels_h), int(M_h), int(n_panels_j), int(M_j))
        st = _lib.Stats()
        h = C.c_void_p()
        if multi:
            kk = np.ascontiguousarray(k0s)
            rc = _lib.lib().qb_window_create_multi(self.ctx.ptr, self.geom.ptr, solver.op,
                                                   self.nwin, kk.ctypes.data, Rs.ctypes.data,
                                                   int(nq), C.byref(cp), _PREC[precision],
                                                   C.byref(h), C.byref(st))
        else:
            rc = _lib.lib().qb_window_create(self.ctx.ptr, self.geom.ptr, solver.op,
                                             self.k0.real, self.k0.imag, self.R, int(nq),
                                             C.byref(cp), _PREC[precision], C.byref(h),
                                             C.byref(st))
        _lib.check(self.ctx.ptr, rc)
        self.ptr = h
        self.stats = st.as_dict()

    def moments(self, V: np.ndarray, refine_tol: float = 1e-13):
        V = np.asfortranarray(V, dtype=np.complex128)
        assert V.shape[0] == self.N
        r = V.shape[1]
        raw0 = np.empty((self.nblocks, r, self.N), dtype=np.complex128)
        raw1 = np.empty((self.nblocks, r, self.N), dtype=np.complex128)
        st = _lib.Stats()
        rc = _lib.lib().qb_window_moments(self.ptr, int(r), V.ctypes.data, raw0.ctypes.data,
                                          raw1.ctypes.data, float(refine_tol), C.byref(st))
        _lib.check(self.ctx.ptr, rc)
        self.stats = st.as_dict()
        if self.nblocks == 1:
            return raw0[0].T, raw1[0].T
        # one (N x r) block per irrep and disk
        return raw0.transpose(0, 2, 1), raw1.transpose(0, 2, 1)

    def moments_qr(self, V: np.ndarray, refine_tol: float = 1e-13):
        """Moments kept on the device and reduced there: A0 = Q R; returns
        (R, C) with C = (Q^H A1)[:r, :], both r x r (qb_window_moments_qr)."""
        V = np.asfortranarray(V, dtype=np.complex128)
        assert V.shape[0] == self.N
        r = V.shape[1]
        R = np.empty((self.nblocks, r, r), dtype=np.complex128)
        Cm = np.empty((self.nblocks, r, r), dtype=np.complex128)
        st = _lib.Stats()
        rc = _lib.lib().qb_window_moments_qr(self.ptr, int(r), V.ctypes.data, R.ctypes.data,
                                             Cm.ctypes.data, float(refine_tol), C.byref(st))
        _lib.check(self.ctx.ptr, rc)
        self.stats = st.as_dict()
        if self.nblocks == 1:
            return R[0].T, Cm[0].T
        # one (R, C) pair per block, stacked: R[b], C[b]
        return R.transpose(0, 2, 1), Cm.transpose(0, 2, 1)

    def basis(self, Ur: np.ndarray, irrep: int = 0) -> np.ndarray:
        """U = Q [Ur; 0] (N x k) from the last moments_qr (qb_window_basis);
        `irrep` selects the moment block (irrep e * nwin + disk c)."""
        Ur = np.asfortranarray(Ur, dtype=np.complex128)
        k = Ur.shape[1]
        U = np.empty((self.N, k), dtype=np.complex128, order="F")
        if k == 0:
            return U
        rc = _lib.lib().qb_window_basis_irrep(self.ptr, int(irrep), Ur.ctypes.data, int(k),
                                              U.ctypes.data)
        _lib.check(self.ctx.ptr, rc)
        return U

    def matrix_cols(self, node: int, j0: int, ncols: int) -> np.ndarray:
        """Columns [j0, j0 + ncols) of the window's own T(z_node), as assembled
        (qb_window_matrix_cols; mixed-precision windows only, an FP64 window
        holds the LU factors).  Returns an N x ncols array."""
        raw = np.empty((int(ncols), self.N), dtype=np.complex128)
        rc = _lib.lib().qb_window_matrix_cols(self.ptr, int(node), int(j0), int(ncols),
                                              raw.ctypes.data)
        _lib.check(self.ctx.ptr, rc)
        return raw.T

    def close(self):
        if getattr(self, "ptr", None):
            _lib.lib().qb_window_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint(N: int, nq: int, precision: str = "fp64", r: int = 64) -> int:
    """Device bytes of one window (the admission footprint of the serving loop)."""
    return int(_lib.lib().qb_window_footprint(int(N), int(nq), _PREC[precision], int(r)))


def _svd_B(A0, A1, svd_tol):
    U, S, Wh = np.linalg.svd(A0, full_matrices=False)
    rk = int(np.count_nonzero(S >= svd_tol))
    return U, S, Wh, rk


def weyl_count(area: float, length: float, k1: float, k2: float) -> float:
    """Dirichlet Weyl estimate N(k2) - N(k1), N(k) = A k^2/(4 pi) - L k/(4 pi)
    (the estimate behind the reference's window planning, beyn.jl:141-160)."""
    Nw = lambda k: area * k * k / (4 * np.pi) - length * k / (4 * np.pi)
    return max(0.0, Nw(k2) - Nw(k1))


def probe_count(billiard, k0: float, R: float, nq: int, svd_tol: float, r0: int = 64) -> int:
    """Probes for a disk: the levels in [k0 - R - d, k0 + R + d] with the
    leak margin d = R (svd_tol^{-1/nq} - 1) of the trapezoid filter, +10 %,
    rounded up to a multiple of r0 (at least r0)."""
    d = R * (svd_tol ** (-1.0 / nq) - 1.0)
    c = weyl_count(billiard.area, billiard.length, k0 - R - d, k0 + R + d)
    r = int(np.ceil((1.1 * c + 16) / r0)) * r0
    return max(r0, r)


def B_from_moments(A0: np.ndarray, A1: np.ndarray, svd_tol: float):
    """Host half of construct_B_matrix (beyn.jl:286-321): thin SVD of A0,
    absolute rank threshold, B = U_k^H A1 W_k Sigma_k^{-1}.  Returns
    (B, Uk, sigma, rk)."""
    U, S, Wh, rk = _svd_B(A0, A1, svd_tol)
    if rk == 0:
        e = np.zeros((A0.shape[0], 0), dtype=np.complex128)
        return np.zeros((0, 0), np.complex128), e, S, 0
    Uk = U[:, :rk]
    Wk = Wh.conj().T[:, :rk]
    B = Uk.conj().T @ ((A1 @ Wk) / S[:rk][None, :])
    return B, Uk, S, rk


def B_from_qr(Rm: np.ndarray, Cm: np.ndarray, svd_tol: float, rel_tol: float = 0.0):
    """B from the device QR of the moments (A0 = Q Rm, Cm = (Q^H A1)[:r]):
    svd(Rm) = U_R S W^H has the singular values and W of svd(A0), and
    B = U_R[:, :k]^H Cm W[:, :k] S_k^{-1} equals U_k^H A1 W_k S_k^{-1}
    (beyn.jl:286-321).  The rank k counts singular values >= svd_tol and
    >= rel_tol sigma_max (rel_tol = 0: the reference's absolute threshold).
    Returns (B, U_R[:, :k], S, rk); U_k = Q U_R[:, :k] comes from
    Window.basis."""
    Ur, S, Wh = np.linalg.svd(Rm)
    thr = max(svd_tol, rel_tol * float(S[0])) if len(S) else svd_tol
    rk = int(np.count_nonzero(S >= thr))
    if rk == 0:
        return np.zeros((0, 0), np.complex128), Ur[:, :0], S, 0
    Wk = Wh.conj().T[:, :rk]
    B = Ur[:, :rk].conj().T @ ((Cm @ Wk) / S[:rk][None, :])
    return B, Ur[:, :rk], S, rk


def construct_B_matrix(solver, pts, N: int, k0: complex, R: float, *, nq: int = 64, r: int = 48,
                       svd_tol: float = 1e-14, rng: Optional[np.random.Generator] = None,
                       use_chebyshev: bool = True, n_panels_h: int = 15000, M_h: int = 5,
                       n_panels_j: int = 3000, M_j: int = 5, info: bool = False,
                       multithreaded: bool = True, precision: str = "fp64",
                       refine_tol: float = 1e-13, grow_r: bool = False,
                       V: Optional[np.ndarray] = None, window: Optional[Window] = None,
                       return_info: bool = False, device_qr: bool = True):
    """beyn.jl:257-322.  Returns (B, Uk) (and a diagnostics dict if
    return_info).  device_qr: reduce the moments on the device (A0 = Q R,
    qb_window_moments_qr) and take the SVD of the r x r R on the host, instead
    of moving the N x r A0, A1 to the host for their SVD; same singular
    values, B and U_k up to rounding."""
    if not use_chebyshev:
        raise RuntimeError("only the Chebyshev path exists for complex contour nodes")
    rng = rng if rng is not None else np.random.default_rng(0)
    assert boundary_matrix_size_for_solver(solver, pts) == N, \
        "N must equal the boundary matrix size"
    t0 = time.perf_counter()
    own = window is None
    win = window or Window(solver, pts, k0, R, nq, n_panels_h=n_panels_h, M_h=M_h,
                           n_panels_j=n_panels_j, M_j=M_j, precision=precision)
    diag = {"window": dict(win.stats)}

    def reduce(Vb):
        # -> (rk, S, finish) with finish() -> (B, Uk)
        if device_qr:
            Rm, Cm = win.moments_qr(Vb, refine_tol)
            th = time.perf_counter()
            B, Ur, S, rk = B_from_qr(Rm, Cm, svd_tol)
            dt = time.perf_counter() - th
            return rk, S, dt, lambda: (B, win.basis(Ur))
        A0, A1 = win.moments(Vb, refine_tol)
        th = time.perf_counter()
        U, S, Wh, rk = _svd_B(A0, A1, svd_tol)
        dt = time.perf_counter() - th

        def finish():
            Uk = U[:, :rk]
            Wk = Wh.conj().T[:, :rk]
            return Uk.conj().T @ ((A1 @ Wk) / S[:rk][None, :]), Uk
        return rk, S, dt, finish

    try:
        Vb = V if V is not None else probe_block(rng, N, r)
        r = Vb.shape[1]
        rk, S, t_svd, finish = reduce(Vb)
        r_used = r
        if rk == 0:
            diag.update(rk=0, r=r_used, sigma=S, seconds=time.perf_counter() - t0)
            e = np.zeros((N, 0), dtype=np.complex128)
            return (e, e, diag) if return_info else (e, e)
        if rk == r:
            r_tmp = r + r
            while r_tmp < N:
                Vb = probe_block(rng, N, r_tmp)
                rk, S, dt, finish = reduce(Vb)
                t_svd += dt
                r_used = r_tmp
                if rk < r_tmp or not grow_r:
                    break
                r_tmp += r_tmp
        th = time.perf_counter()
        B, Uk = finish()
        t_svd += time.perf_counter() - th
        diag.update(rk=rk, r=r_used, sigma=S, moments=dict(win.stats), t_host_svd_s=t_svd,
                    seconds=time.perf_counter() - t0)
        return (B, Uk, diag) if return_info else (B, Uk)
    finally:
        if own:
            win.close()


def solve_vect(solver, basis, pts, k: complex, dk: float, *, nq: int = 32, r: int = 48,
               svd_tol: float = 1e-14, res_tol: float = 1e-8, rng=None, **kw):
    """beyn.jl:364-375 -> (lam, Uk, Y, k, dk, pts)."""
    N = boundary_matrix_size_for_solver(solver, pts)
    B, Uk = construct_B_matrix(solver, pts, N, k, dk, nq=nq, r=r, svd_tol=svd_tol, rng=rng, **kw)
    if B.size == 0:
        return np.zeros(0, np.complex128), Uk, np.zeros((0, 0), np.complex128), k, dk, pts
    lam, Y = np.linalg.eig(B)
    return lam, Uk, Y, k, dk, pts


def solve(solver, basis, pts, k: complex, dk: float, *, nq: int = 32, r: int = 48,
          svd_tol: float = 1e-14, res_tol: float = 1e-8, rng=None, **kw):
    """beyn.jl:386-397 -> lam (no filtering, like the reference)."""
    N = boundary_matrix_size_for_solver(solver, pts)
    B, Uk = construct_B_matrix(solver, pts, N, k, dk, nq=nq, r=r, svd_tol=svd_tol, rng=rng, **kw)
    if B.size == 0:
        return np.zeros(0, np.complex128)
    return np.linalg.eigvals(B)
