"""Spectrum driver around the Beyn windows: Weyl window planning, the
residual check of Beyn eigenpairs and compute_spectrum_beyn.

Mirrors (src/solvers/acceleratedmethods/beyn.jl):
  delta_weyl / initial_step_from_dos / grow_upper_bound / bisect_for_delta_k
                                                               beyn.jl:60-135
  plan_weyl_windows(billiard, k1, k2; m, Rmax, fundamental, tol_levels, maxit)
                                                               beyn.jl:141-160
  beyn_disks_from_windows(iv) -> (k0, R)                       beyn.jl:167-173
  residual_and_norm_select(solver, lam, Uk, Y, k0, R, pts; res_tol, matnorm,
                           epss, auto_discard_spurious, collect_logs, ...)
      -> (idx, Phi_kept, tens, tensN, logs)                    beyn.jl:557-593
  imag_k_check_EXPERIMENTAL(solver, lams, Uks, Ys, k0s, Rs, all_pts; res_tol,
                            pad, group_size, ...) -> (idx_keep, residuals)
                                                               beyn.jl:623-721
  compute_spectrum_beyn(solver, billiard, k1, k2; m, Rmax, nq, r, svd_tol,
                        res_tol, ...) -> (ks, tens, us, pts, tensN)
                                                               beyn.jl:778-921
The residuals ||T(lambda) phi|| are computed on the device (qb_residuals):
T(lambda) is assembled from the same Chebyshev tables as the contour matrices
and applied to phi = U_k Y[:, j] without leaving the device; only the norms
come back.  Indices are 0-based where the reference's are 1-based.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .beyn import construct_B_matrix
from .context import Context, default_context
from .geometry import Billiard, evaluate_points
from .solvers import (Geometry, _cheb, boundary_matrix_size, boundary_matrix_size_for_solver,
                      chebyshev_params)

EPS = np.finfo(np.float64).eps


# -- Weyl window planning (beyn.jl:60-173) -----------------------------------
def _area(billiard: Billiard, fundamental: bool) -> float:
    if fundamental and getattr(billiard, "area_fundamental", None) is not None:
        return billiard.area_fundamental
    return billiard.area


def delta_weyl(billiard: Billiard, k: float, dk: float, fundamental: bool = True) -> float:
    """Area term of Weyl's law, A((k + dk)^2 - k^2)/(4 pi) (beyn.jl:71-74, :80-83)."""
    return _area(billiard, fundamental) * ((k + dk) ** 2 - k * k) / (4 * math.pi)


def initial_step_from_dos(billiard, k, m, fundamental=True, min_step=1e-6):
    """m / rho(k) with rho = A k / 2pi, at least min_step (beyn.jl:85-88)."""
    rho = max(_area(billiard, fundamental) * k / (2 * math.pi), 1e-12)
    return max(m / rho, min_step)


def grow_upper_bound(billiard, k, m, dk0, remaining, fundamental=True, max_grows=60):
    """Double dk until the window holds >= m levels or reaches `remaining`
    (beyn.jl:90-98).  Returns (dk, reached)."""
    dk = min(remaining, max(dk0, np.spacing(k)))
    for _ in range(max_grows):
        if delta_weyl(billiard, k, dk, fundamental) >= m - EPS:
            return dk, True
        if dk >= 0.999999 * remaining:
            return remaining, False
        dk = min(remaining, 2 * dk)
    return dk, delta_weyl(billiard, k, dk, fundamental) >= m


def bisect_for_delta_k(billiard, k, m, lo, hi, fundamental=True, tol_levels=0.1, maxit=50):
    """Bisection on dk for |N(dk) - m| <= tol_levels (beyn.jl:100-115)."""
    assert hi > lo
    Nlo = delta_weyl(billiard, k, lo, fundamental)
    if abs(Nlo - m) <= tol_levels:
        return lo
    Nhi = delta_weyl(billiard, k, hi, fundamental)
    if abs(Nhi - m) <= tol_levels:
        return hi
    assert Nhi >= m
    for _ in range(maxit):
        mid = 0.5 * (lo + hi)
        Nmid = delta_weyl(billiard, k, mid, fundamental)
        if abs(Nmid - m) <= tol_levels:
            return mid
        if Nmid < m:
            lo = mid
        else:
            hi = mid
        if hi - lo <= max(1e-12, 1e-9 * max(1.0, k)):
            return 0.5 * (lo + hi)
    return 0.5 * (lo + hi)


def plan_weyl_windows(billiard: Billiard, k1: float, k2: float, *, m: int = 10,
                      Rmax: float = 1.0, fundamental: bool = True, tol_levels: float = 0.1,
                      maxit: int = 50) -> List[tuple]:
    """Cover [k1, k2] with windows of ~m Weyl levels, each at most 2 Rmax long
    (beyn.jl:141-160)."""
    iv = []
    k = k1
    while k < k2 - EPS:
        rem_raw = k2 - k
        rem = 2 * Rmax if rem_raw > 2 * Rmax else rem_raw
        if delta_weyl(billiard, k, rem, fundamental) <= m + tol_levels and rem == rem_raw:
            iv.append((k, k2))
            break
        dk0 = initial_step_from_dos(billiard, k, m, fundamental)
        hi, ok = grow_upper_bound(billiard, k, m, dk0, rem, fundamental)
        if not ok:
            iv.append((k, k + hi))
            k += hi
            continue
        dk = bisect_for_delta_k(billiard, k, m, 0.0, hi, fundamental, tol_levels, maxit)
        kR = min(k + dk, k + rem)
        iv.append((k, kR))
        k = kR
    return iv


def beyn_disks_from_windows(iv: Sequence[tuple]):
    """Centres (complex, imaginary part 0) and radii of the windows
    (beyn.jl:167-173)."""
    k0 = np.array([complex(0.5 * (a + b)) for a, b in iv], dtype=np.complex128)
    R = np.array([0.5 * (b - a) for a, b in iv], dtype=np.float64)
    return k0, R


# -- residuals on the device ---------------------------------------------------
def residual_norms(solver, pts, lam, Phi, *, n_panels_h=15000, M_h=5, n_panels_j=3000, M_j=5,
                   geom: Optional[Geometry] = None, ctx: Optional[Context] = None,
                   batch: int = 0):
    """(||y_q||_2, ||y_q||_1, ||y_q||_inf, ||T(lam_q)||_1, ||T(lam_q)||_inf)
    with y_q = T(lam_q) Phi[:, q], for every q (qb_residuals)."""
    ctx = ctx or (geom.ctx if geom is not None else default_context())
    g = geom or Geometry(pts, ctx, solver)
    lam = np.ascontiguousarray(np.asarray(lam, dtype=np.complex128).ravel())
    nl = len(lam)
    Phi = np.asfortranarray(Phi, dtype=np.complex128)
    assert Phi.shape == (g.dim, nl), f"Phi must be {g.dim} x {nl}"
    outs = [np.zeros(nl) for _ in range(5)]
    if nl == 0:
        return tuple(outs)
    cp = _cheb(n_panels_h, M_h, n_panels_j, M_j)
    rc = _lib.lib().qb_residuals(ctx.ptr, g.ptr, solver.op, lam.ctypes.data, nl, C.byref(cp),
                                 Phi.ctypes.data, int(batch), *[o.ctypes.data for o in outs])
    _lib.check(ctx.ptr, rc)
    return tuple(outs)


def residual_and_norm_select(solver, lam, Uk, Y, k0, R, pts, *, res_tol: float,
                             matnorm: str = "one", epss: float = 1e-15,
                             auto_discard_spurious: bool = True, collect_logs: bool = False,
                             use_chebyshev: bool = True, n_panels_h: int = 15000, M_h: int = 5,
                             n_panels_j: int = 3000, M_j: int = 5, multithreaded: bool = True,
                             geom: Optional[Geometry] = None, ctx: Optional[Context] = None):
    """beyn.jl:557-593: keep the eigenpairs inside the disk whose residual
    ||T(lam_j) phi_j||_2, phi_j = Uk Y[:, j], is below res_tol (all inside
    ones if not auto_discard_spurious).  Returns (idx, Phi_kept, tens[idx],
    tensN[idx], logs) with tensN = ||y|| / (||T|| (||phi|| + epss) + epss) in
    the `matnorm` norm ("one" or "inf"; the operator 2-norm is not provided)."""
    if not use_chebyshev:
        raise RuntimeError("only the Chebyshev path exists for complex eigenvalues")
    if matnorm not in ("one", "inf"):
        raise NotImplementedError("matnorm must be 'one' or 'inf'")
    lam = np.asarray(lam, dtype=np.complex128)
    N, rk = Uk.shape
    inside = [j for j in range(len(lam)) if abs(lam[j] - k0) <= R]
    tens = np.full(rk, np.nan)
    tensN = np.full(rk, np.nan)
    Phi = Uk @ Y if rk else np.zeros((N, 0), np.complex128)
    logs = []
    keep = np.zeros(rk, bool)
    if inside:
        r2, r1, rinf, n1, ninf = residual_norms(solver, pts, lam[inside], Phi[:, inside],
                                          n_panels_h=n_panels_h, M_h=M_h, n_panels_j=n_panels_j,
                                          M_j=M_j, geom=geom, ctx=ctx)
        for q, j in enumerate(inside):
            if matnorm == "one":
                nA, phin, yn = n1[q], float(np.sum(np.abs(Phi[:, j]))), r1[q]
            else:
                nA, phin, yn = ninf[q], float(np.max(np.abs(Phi[:, j]))), rinf[q]
            tens[j] = r2[q]
            tensN[j] = yn / (nA * (phin + epss) + epss)
            if auto_discard_spurious and r2[q] >= res_tol:
                if collect_logs:
                    logs.append(f"λ={lam[j]} ||Aφ||={r2[q]} > {res_tol} → DROP")
            else:
                keep[j] = True
                if collect_logs:
                    logs.append(f"λ={lam[j]} ||Aφ||={r2[q]} < {res_tol} ← KEEP")
    idx = np.flatnonzero(keep)
    Phi_kept = Phi[:, idx] if len(idx) else np.zeros((N, 0), np.complex128)
    return idx, Phi_kept, tens[idx], tensN[idx], logs


def imag_k_check_EXPERIMENTAL(solver, lams, Uks, Ys, k0s, Rs, all_pts, *, res_tol: float,
                              pad: int = 20, group_size: int = 100, use_chebyshev: bool = True,
                              n_panels_h: int = 15000, M_h: int = 5, n_panels_j: int = 3000,
                              M_j: int = 5, multithreaded: bool = True, verbose: bool = False,
                              geoms: Optional[Sequence[Geometry]] = None,
                              ctx: Optional[Context] = None):
    """beyn.jl:623-721: residual-check the inside candidates of all windows in
    order of decreasing |Im lambda|, groups of `group_size`, drop those with
    residual >= res_tol, stop after `pad` consecutive good ones.  Returns
    (idx_keep, residuals) per window; residuals are NaN for unchecked roots."""
    if not use_chebyshev:
        raise RuntimeError("only the Chebyshev path exists for complex eigenvalues")
    nw = len(lams)
    idx_inside, residuals, cand = [], [], []
    local_pos = {}
    for i in range(nw):
        li = np.asarray(lams[i])
        ins = [j for j in range(len(li)) if abs(li[j] - k0s[i]) <= Rs[i]]
        idx_inside.append(ins)
        residuals.append(np.full(len(ins), np.nan))
        for lp, j in enumerate(ins):
            local_pos[(i, j)] = lp
            cand.append((i, j, abs(li[j].imag), li[j].real))
    cand.sort(key=lambda c: c[2], reverse=True)  # stable, like Julia's sort!
    drop = set()
    checked = dropped = streak = 0
    pos, stop_early = 0, False
    while pos < len(cand) and not stop_early:
        group = cand[pos:pos + group_size]
        rdict = {}
        for iwin in dict.fromkeys(c[0] for c in group):
            sub = [c for c in group if c[0] == iwin]
            lam_g = np.array([lams[i][j] for (i, j, _, _) in sub])
            Phi = np.column_stack([Uks[i] @ Ys[i][:, j] for (i, j, _, _) in sub])
            g = geoms[iwin] if geoms is not None else None
            r2 = residual_norms(solver, all_pts[iwin], lam_g, Phi, n_panels_h=n_panels_h,
                                         M_h=M_h, n_panels_j=n_panels_j, M_j=M_j, geom=g,
                                         ctx=ctx)[0]
            for q, (i, j, _, _) in enumerate(sub):
                rdict[(i, j)] = r2[q]
        for (i, j, imj, _) in group:
            rj = rdict[(i, j)]
            checked += 1
            residuals[i][local_pos[(i, j)]] = rj
            if rj >= res_tol:
                drop.add((i, j))
                dropped += 1
                streak = 0
            else:
                streak += 1
                if streak >= pad:
                    stop_early = True
                    break
        pos += group_size
    idx_keep = []
    for i in range(nw):
        mask = np.array([(i, j) not in drop for j in idx_inside[i]], dtype=bool)
        idx_keep.append(np.asarray(idx_inside[i], dtype=int)[mask] if len(mask) else
                        np.zeros(0, int))
        residuals[i] = residuals[i][mask] if len(mask) else residuals[i]
    if verbose:
        print(f"grouped tail residual check: checked={checked} dropped={dropped} "
              f"total_candidates={len(cand)}")
    return idx_keep, residuals


# -- the driver (beyn.jl:778-921) --------------------------------------------
def compute_spectrum_beyn(solver, billiard: Billiard, k1: float, k2: float, *, m: int = 50,
                          Rmax: float = 0.5, nq: int = 48, r: Optional[int] = None,
                          svd_tol: float = 1e-12, res_tol: float = 1e-9,
                          auto_discard_spurious: bool = True, multithreaded_matrix: bool = True,
                          use_adaptive_svd_tol: bool = False, use_chebyshev: bool = True,
                          n_panels_h: int = 15000, M_h: int = 5, n_panels_j: int = 3000,
                          M_j: int = 5, do_INFO_init: bool = False,
                          do_per_solve_INFO: bool = False, cheb_tol: float = 1e-13,
                          max_iter: int = 20, sampling_points: int = 50_000,
                          grow_panels: float = 1.5, grow_M: int = 2,
                          return_imag_part: bool = False,
                          use_imag_check_EXPERIMENTAL: bool = True, precision: str = "fp64",
                          refine_tol: float = 1e-13, grow_r: bool = False,
                          ctx: Optional[Context] = None):
    """Weyl-planned Beyn sweep of [k1, k2]: plan disks of ~m levels, one
    discretisation per disk at its centre, one Chebyshev table size for all
    disks (tuned on the disk reaching highest k), a Beyn solve per disk, and
    the spurious-root filter.  Returns (ks, tens, us, pts, tensN) in window
    order, as the reference does.  The per-solve INFO diagnostics
    (solve_INFO) are not reproduced."""
    if grow_panels != 1.5 or grow_M != 2:
        raise NotImplementedError("the table tuner grows panels x1.5 and M by 2")
    if use_adaptive_svd_tol:
        raise NotImplementedError("use_adaptive_svd_tol is not provided")
    r = m + 15 if r is None else r
    ctx = ctx or default_context()
    fundamental = solver.symmetry is not None
    if fundamental and billiard.area_fundamental is None:
        # the leading Weyl term of one symmetry sector
        billiard = dataclasses.replace(billiard,
                                       area_fundamental=billiard.area / solver.symmetry.order)
    intervals = plan_weyl_windows(billiard, k1, k2, m=m, fundamental=fundamental, Rmax=Rmax)
    if len(intervals) >= 2:
        kL2, kR2 = intervals[-2]
        kL3, kR3 = intervals[-1]
        if kR3 - kL3 <= max(100 * np.spacing(k2), 1e-9 * max(k2, 1.0)):
            if kR3 - kL2 <= 2 * Rmax + 10 * np.spacing(k2):
                intervals[-2] = (kL2, kR3)
            intervals.pop()
    k0, R = beyn_disks_from_windows(intervals)
    if len(k0) == 0:
        return np.zeros(0), np.zeros(0), [], [], np.zeros(0)
    if nq <= 15:
        raise ValueError("Do not use less than 15 contour nodes")
    all_pts = [evaluate_points(solver, billiard, float(k.real)) for k in k0]
    geoms = [Geometry(p, ctx, solver) for p in all_pts]
    try:
        if use_chebyshev:
            imax = int(np.argmax(k0.real + R))
            th = np.linspace(0.0, 2 * np.pi, nq + 1)[:-1]
            zj_ref = k0[imax] + R[imax] * np.exp(1j * th)
            n_panels_h, M_h, n_panels_j, M_j, _ = chebyshev_params(
                solver, all_pts[imax], zj_ref, npanels_h_init=n_panels_h, M_h_init=M_h,
                npanels_j_init=n_panels_j, M_j_init=M_j, tol=cheb_tol,
                sampling_points=sampling_points, max_iter=max_iter, geom=geoms[imax])
        cheb = dict(n_panels_h=n_panels_h, M_h=M_h, n_panels_j=n_panels_j, M_j=M_j)
        lams, Uks, Ys = [], [], []
        for i in range(len(k0)):
            B, Uk = _window_B(solver, all_pts[i], geoms[i], complex(k0[i]), float(R[i]), nq, r,
                              svd_tol, precision, refine_tol, grow_r, cheb)
            if B.size == 0:
                lams.append(np.zeros(0, np.complex128))
                Uks.append(Uk)
                Ys.append(np.zeros((0, 0), np.complex128))
                continue
            lam, Y = np.linalg.eig(B)
            lams.append(lam)
            Uks.append(Uk)
            Ys.append(Y)
        ks_l, tens_l, tensN_l, phi_l = [], [], [], []
        if use_imag_check_EXPERIMENTAL:
            idx_keep, resid = imag_k_check_EXPERIMENTAL(
                solver, lams, Uks, Ys, list(k0), list(R), all_pts, res_tol=res_tol, pad=20,
                group_size=100, geoms=geoms, ctx=ctx, verbose=do_per_solve_INFO, **cheb)
            for i in range(len(k0)):
                idx = idx_keep[i]
                li = lams[i]
                ks_l.append(li[idx] if return_imag_part else li[idx].real)
                tens_l.append(np.abs(li[idx].imag))
                tensN_l.append(resid[i])
                phi_l.append(Uks[i] @ Ys[i][:, idx] if len(idx) else
                             np.zeros((boundary_matrix_size_for_solver(solver, all_pts[i]), 0), np.complex128))
        else:
            for i in range(len(k0)):
                if len(lams[i]) == 0:
                    ks_l.append(np.zeros(0, np.complex128 if return_imag_part else np.float64))
                    tens_l.append(np.zeros(0))
                    tensN_l.append(np.zeros(0))
                    phi_l.append(np.zeros((boundary_matrix_size_for_solver(solver, all_pts[i]), 0), np.complex128))
                    continue
                idx, Phi, traw, tnorm, _ = residual_and_norm_select(
                    solver, lams[i], Uks[i], Ys[i], k0[i], R[i], all_pts[i], res_tol=res_tol,
                    matnorm="one", epss=1e-15, auto_discard_spurious=auto_discard_spurious,
                    geom=geoms[i], ctx=ctx, **cheb)
                li = lams[i]
                ks_l.append(li[idx] if return_imag_part else li[idx].real)
                tens_l.append(traw)
                tensN_l.append(tnorm)
                phi_l.append(Phi)
        ks = np.concatenate(ks_l) if ks_l else np.zeros(0)
        tens = np.concatenate(tens_l) if tens_l else np.zeros(0)
        tensN = np.concatenate(tensN_l) if tensN_l else np.zeros(0)
        us, pts_all = [], []
        for i, Phi in enumerate(phi_l):
            for j in range(Phi.shape[1]):
                us.append(Phi[:, j].copy())
                pts_all.append(all_pts[i])
        return ks, tens, us, pts_all, tensN
    finally:
        for g in geoms:
            g.close()


def _window_B(solver, pts, geom, k0, R, nq, r, svd_tol, precision, refine_tol, grow_r, cheb):
    """construct_B_matrix on a resident geometry (one window, rng seeded 0
    per disk like the reference's MersenneTwister(0))."""
    from .beyn import Window
    win = Window(solver, pts, k0, R, nq, precision=precision, geom=geom, **cheb)
    try:
        return construct_B_matrix(solver, pts, geom.dim, k0, R, nq=nq, r=r, svd_tol=svd_tol,
                                  rng=np.random.default_rng(0), precision=precision,
                                  refine_tol=refine_tol, grow_r=grow_r, window=win, **cheb)
    finally:
        win.close()
