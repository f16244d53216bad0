"""Kress solver structs and construct_boundary_matrices! on libqbhip.

Mirrors the reference's operator boundary for the two Kress families:
  construct_boundary_matrices!(Tbufs, solver, pts, zj; multithreaded, use_chebyshev,
                               n_panels_h=15000, M_h=5, n_panels_j=3000, M_j=5, timeit)
  * CFIE_kress: src/chebyshev/chebyshev_euclidian_helmholtz_cfie_kress.jl:1431-1450
  * DLP_kress:  src/chebyshev/chebyshev_euclidian_helmholtz_dlp_kress.jl:1110-1124
Tbufs[q] is overwritten with T(zj[q]) (column-major N x N complex128, the
layout of Julia's Matrix{ComplexF64}).  Error behaviour follows the
reference: a length/size mismatch of Tbufs raises AssertionError
(cheb-CFIE:1433, cheb-DLP:1117); use_chebyshev=false raises for DLP
(cheb-DLP:1111) and, where the reference logs @error and leaves zeros for
CFIE (cheb-CFIE:1447), this mirror raises as well rather than return zeros.
"""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _lib
from .context import Context, default_context
from .geometry import Billiard, BoundaryPointsCFIE
from .symmetry import orbits_for


@dataclass
class CFIE_kress:
    """CFIE_kress(pts_scaling_factor, billiard; min_pts=20, symmetry=nothing)
    (src/solvers/sweepmethods/cfie/cfie.jl:92-101, :266-271)."""
    pts_scaling_factor: float
    billiard: Billiard
    min_pts: int = 20
    symmetry: object = None
    op: int = field(default=_lib.QB_CFIE_KRESS, init=False)


@dataclass
class DLP_kress:
    """DLP_kress(pts_scaling_factor, billiard; min_pts, symmetry)
    (src/solvers/sweepmethods/dlp/dlp_kress.jl:137-146, :215-220)."""
    pts_scaling_factor: float
    billiard: Billiard
    min_pts: int = 20
    symmetry: object = None
    op: int = field(default=_lib.QB_DLP_KRESS, init=False)


def boundary_matrix_size(pts) -> int:
    """boundary_matrix_size / component_offsets (boundarypoints.jl:549-576)."""
    if isinstance(pts, BoundaryPointsCFIE):
        return len(pts)
    return int(sum(len(p) for p in pts))


def boundary_matrix_size_for_solver(solver, pts) -> int:
    """boundary_matrix_size_for_solver (ebim.jl:26-40): the number of
    symmetry orbits when the solver carries a symmetry, else the node count."""
    sym = getattr(solver, "symmetry", None)
    if sym is None:
        return boundary_matrix_size(pts)
    return boundary_matrix_size(pts) // sym.order


def _as_list(pts) -> List[BoundaryPointsCFIE]:
    return [pts] if isinstance(pts, BoundaryPointsCFIE) else list(pts)


class Geometry:
    """A boundary discretisation resident on the device (qb_geom).  With a
    solver that carries a symmetry, the geometry also holds the solver's
    orbit table (qb_geom_set_orbits) and every assembly on it is the
    symmetry-reduced one; `N` is the node count, `dim` the matrix dimension."""

    def __init__(self, pts, ctx: Optional[Context] = None, solver=None, irreps=None):
        """irreps: solvers with symmetries of one group (e.g. Rotation(5, m),
        m = 0..4), assembled together (qb_geom_set_orbits_multi): every
        assembly or window on this geometry then covers all of them."""
        self.ctx = ctx or default_context()
        comps = _as_list(pts)
        cat = lambda f: np.ascontiguousarray(np.concatenate([f(p) for p in comps]), dtype=np.float64)
        self._arrs = [cat(lambda p: p.xy[:, 0]), cat(lambda p: p.xy[:, 1]),
                      cat(lambda p: p.tangent[:, 0]), cat(lambda p: p.tangent[:, 1]),
                      cat(lambda p: p.tangent_2[:, 0]), cat(lambda p: p.tangent_2[:, 1]),
                      cat(lambda p: p.ts), cat(lambda p: p.ws)]
        self.n_per_comp = np.array([len(p) for p in comps], dtype=np.int64)
        self.ncomp = len(comps)
        self.N = int(self.n_per_comp.sum())
        L = _lib.lib()
        h = C.c_void_p()
        rc = L.qb_geom_create(self.ctx.ptr, self.ncomp, self.n_per_comp.ctypes.data,
                              *[a.ctypes.data for a in self._arrs], C.byref(h))
        _lib.check(self.ctx.ptr, rc)
        self.ptr = h
        self.dim = self.N
        self.nirr = 1
        self.symmetry = getattr(solver, "symmetry", None)
        orb = orbits_for(solver, pts) if solver is not None else None
        if irreps:
            tabs = [orbits_for(s_, pts) for s_ in irreps]
            img = np.ascontiguousarray(tabs[0][0], dtype=np.int64)
            if any(t is None or not np.array_equal(t[0], img) for t in tabs):
                raise ValueError("irreps must share one orbit table (one symmetry group)")
            sc = np.ascontiguousarray(np.stack([t[1] for t in tabs]), dtype=np.complex128)
            rc = L.qb_geom_set_orbits_multi(self.ctx.ptr, self.ptr, img.shape[0], img.shape[1],
                                            len(tabs), img.ctypes.data, sc.ctypes.data)
            _lib.check(self.ctx.ptr, rc)
            self.dim = int(img.shape[0])
            self.nirr = len(tabs)
            self.irreps = list(irreps)
            self.orbits = (img, sc)
            orb = None
        if orb is not None:
            img, sc = orb
            img = np.ascontiguousarray(img, dtype=np.int64)
            sc = np.ascontiguousarray(sc, dtype=np.complex128)
            rc = L.qb_geom_set_orbits(self.ctx.ptr, self.ptr, img.shape[0], img.shape[1],
                                      img.ctypes.data, sc.ctypes.data)
            _lib.check(self.ctx.ptr, rc)
            self.dim = int(img.shape[0])
            self.orbits = (img, sc)

    def close(self):
        if getattr(self, "ptr", None):
            _lib.lib().qb_geom_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _cheb(n_panels_h, M_h, n_panels_j, M_j):
    return _lib.ChebParams(int(n_panels_h), int(M_h), int(n_panels_j), int(M_j))


def assemble(solver, pts, zj, *, n_panels_h=15000, M_h=5, n_panels_j=3000, M_j=5,
             geom: Optional[Geometry] = None, ctx: Optional[Context] = None,
             out: Optional[np.ndarray] = None, conj_pairs: bool = False) -> np.ndarray:
    """All T(z_q) in one launch.  Returns an array `T` of shape (nz, N, N)
    with T[q] the matrix T(z_q) in the usual (row, column) indexing; its
    storage is column-major per matrix (T[q].flags.f_contiguous).  With a
    solver symmetry, N is the number of orbits (the reduced matrices).
    conj_pairs=True takes the kernel of Beyn windows with real k0
    (QB_ASM_CONJ_PAIRS): zj must be closed under conjugation, and both
    matrices of a pair come from the two nodes' Hankel tables."""
    ctx = ctx or (geom.ctx if geom is not None else default_context())
    g = geom or Geometry(pts, ctx, solver)
    zj = np.ascontiguousarray(np.asarray(zj, dtype=np.complex128).ravel())
    nz, N = len(zj), g.dim
    nm = nz * getattr(g, "nirr", 1)  # matrix e * nz + q for irrep e of a multi-irrep geometry
    if out is None:
        raw = np.empty((nm, N, N), dtype=np.complex128)
    else:
        raw = out
        assert raw.shape == (nm, N, N) and raw.dtype == np.complex128 and raw.flags.c_contiguous
    cp = _cheb(n_panels_h, M_h, n_panels_j, M_j)
    rr = np.zeros(2)
    flags = _lib.QB_ASM_CONJ_PAIRS if conj_pairs else 0
    rc = _lib.lib().qb_assemble(ctx.ptr, g.ptr, solver.op, zj.ctypes.data, nz, C.byref(cp),
                                raw.ctypes.data, flags, rr.ctypes.data)
    _lib.check(ctx.ptr, rc)
    # raw[q] holds T(z_q) column-major: raw[q, j, i] = T[i, j]
    return raw.transpose(0, 2, 1)


def construct_boundary_matrices(Tbufs, solver, pts, zj, *, multithreaded: bool = True,
                                use_chebyshev: bool = True, n_panels_h: int = 15000,
                                M_h: int = 5, n_panels_j: int = 3000, M_j: int = 5,
                                timeit: bool = False, geom: Optional[Geometry] = None,
                                ctx: Optional[Context] = None):
    """construct_boundary_matrices!(Tbufs, solver, pts, zj; ...) (see module
    docstring).  Tbufs: sequence of (N, N) complex128 arrays, overwritten."""
    if not use_chebyshev:
        raise RuntimeError("Direct complex-k construction is not implemented; "
                           "use_chebyshev=true is the only complex-k path")
    zj = np.asarray(zj, dtype=np.complex128).ravel()
    assert len(Tbufs) == len(zj), "length(Tbufs) must equal length(zj)"
    if isinstance(solver, DLP_kress) and not isinstance(pts, BoundaryPointsCFIE):
        comps = _as_list(pts)
        assert len(comps) == 1, "DLP_kress expects one BoundaryPointsCFIE component"
    N = boundary_matrix_size_for_solver(solver, pts)
    for q, T in enumerate(Tbufs):
        assert T.shape == (N, N), f"Tbufs[{q}] has size {T.shape}, but the workspace requires ({N},{N})."
    T = assemble(solver, pts, zj, n_panels_h=n_panels_h, M_h=M_h, n_panels_j=n_panels_j,
                 M_j=M_j, geom=geom, ctx=ctx)
    for q in range(len(zj)):
        np.copyto(Tbufs[q], T[q])
    return None


def assemble_with_derivatives(solver, pts, zj, *, n_panels_h=15000, M_h=5, n_panels_j=3000,
                              M_j=5, geom: Optional[Geometry] = None,
                              ctx: Optional[Context] = None,
                              reference_offblock_sign: bool = True):
    """T(z_q), dT/dk(z_q) and d2T/dk2(z_q) in one launch (qb_assemble_deriv).
    Returns three arrays of shape (nz, N, N) indexed like `assemble`.
    Between boundary components the two double-layer derivative terms carry
    the reference's sign (cheb-CFIE:1031-1032) by default, so the drop-in
    returns the reference's numbers; reference_offblock_sign=False gives them
    the opposite sign, which makes dT/dk the derivative of T there."""
    ctx = ctx or (geom.ctx if geom is not None else default_context())
    g = geom or Geometry(pts, ctx, solver)
    zj = np.ascontiguousarray(np.asarray(zj, dtype=np.complex128).ravel())
    nz, N = len(zj), g.dim
    raws = [np.empty((nz, N, N), dtype=np.complex128) for _ in range(3)]
    cp = _cheb(n_panels_h, M_h, n_panels_j, M_j)
    rr = np.zeros(2)
    flags = _lib.QB_DERIV_REFERENCE_SIGN if reference_offblock_sign else _lib.QB_DERIV_MATRIX_SIGN
    rc = _lib.lib().qb_assemble_deriv(ctx.ptr, g.ptr, solver.op, zj.ctypes.data, nz, C.byref(cp),
                                      *[r.ctypes.data for r in raws], flags, rr.ctypes.data)
    _lib.check(ctx.ptr, rc)
    return tuple(r.transpose(0, 2, 1) for r in raws)


def construct_boundary_matrices_with_derivatives(Tbufs, dTbufs, ddTbufs, solver, pts, zj, *,
                                                 multithreaded: bool = True,
                                                 use_chebyshev: bool = True,
                                                 n_panels_h: int = 15000, M_h: int = 5,
                                                 n_panels_j: int = 3000, M_j: int = 5,
                                                 timeit: bool = False,
                                                 geom: Optional[Geometry] = None,
                                                 ctx: Optional[Context] = None,
                                                 reference_offblock_sign: bool = True):
    """construct_boundary_matrices_with_derivatives!(Tbufs, dTbufs, ddTbufs,
    solver, pts, zj; ...) (cheb-CFIE:1507-1530, cheb-DLP:1126-1144): the
    three buffers of every wavenumber are overwritten with T, dT/dk and
    d2T/dk2."""
    if not use_chebyshev:
        raise RuntimeError("Direct derivative matrix construction is only for real k")
    zj = np.asarray(zj, dtype=np.complex128).ravel()
    Mk = len(zj)
    assert len(Tbufs) == Mk and len(dTbufs) == Mk and len(ddTbufs) == Mk
    if isinstance(solver, DLP_kress) and not isinstance(pts, BoundaryPointsCFIE):
        assert len(_as_list(pts)) == 1, "DLP_kress expects one BoundaryPointsCFIE component"
    N = boundary_matrix_size_for_solver(solver, pts)
    for bufs in (Tbufs, dTbufs, ddTbufs):
        for q, T in enumerate(bufs):
            assert T.shape == (N, N), f"buffer {q} has size {T.shape}, but ({N},{N}) is required"
    A, A1, A2 = assemble_with_derivatives(solver, pts, zj, n_panels_h=n_panels_h, M_h=M_h,
                                          n_panels_j=n_panels_j, M_j=M_j, geom=geom, ctx=ctx,
                                          reference_offblock_sign=reference_offblock_sign)
    for q in range(Mk):
        np.copyto(Tbufs[q], A[q])
        np.copyto(dTbufs[q], A1[q])
        np.copyto(ddTbufs[q], A2[q])
    return None


def chebyshev_params(solver, pts, zj, *, npanels_h_init=15000, M_h_init=5, npanels_j_init=3000,
                     M_j_init=5, tol=1e-10, sampling_points=50000, max_iter=20,
                     geom: Optional[Geometry] = None, ctx: Optional[Context] = None):
    """chebyshev_params (chebyshev_optimal_panelization.jl:221-262) on the
    device: returns (nh, Mh, nj, Mj, errs[4])."""
    ctx = ctx or (geom.ctx if geom is not None else default_context())
    g = geom or Geometry(pts, ctx)
    zj = np.ascontiguousarray(np.asarray(zj, dtype=np.complex128).ravel())
    cp = _cheb(npanels_h_init, M_h_init, npanels_j_init, M_j_init)
    errs = np.zeros(4)
    rc = _lib.lib().qb_cheb_tune(ctx.ptr, g.ptr, solver.op, zj.ctypes.data, len(zj), float(tol),
                                 int(sampling_points), int(max_iter), C.byref(cp),
                                 errs.ctypes.data)
    _lib.check(ctx.ptr, rc)
    if not np.all(errs < tol):
        # the reference warns and returns its last tables as well
        # (chebyshev_optimal_panelization.jl:260-261)
        warnings.warn(f"chebyshev_params: tolerance {tol:g} not met after {max_iter} iterations "
                      f"(errors H0 {errs[0]:.3g}, H1 {errs[1]:.3g}, J0 {errs[2]:.3g}, "
                      f"J1 {errs[3]:.3g}); returning the last table sizes",
                      RuntimeWarning, stacklevel=2)
    return cp.n_panels_h, cp.M_h, cp.n_panels_j, cp.M_j, errs


def bessel_eval(z, ctx: Optional[Context] = None):
    """Device H0^(1), H1^(1), J0, J1 at complex points z."""
    ctx = ctx or default_context()
    z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128).ravel())
    outs = [np.empty_like(z) for _ in range(4)]
    rc = _lib.lib().qb_bessel_eval(ctx.ptr, z.ctypes.data, len(z), *[o.ctypes.data for o in outs])
    _lib.check(ctx.ptr, rc)
    return tuple(outs)
