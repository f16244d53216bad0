"""Continuous-batching server for Beyn spectral windows.

A request is one disk (geometry, k0, contour radius R, tolerance).  The
server keeps the boundary quadrature of every (shape, N) resident on the
device and runs windows in `slots` device contexts (each with its own HIP
streams), so windows admitted together overlap on the GPU.  A request is
admitted when a slot is free and its HBM footprint (the N^2 matrices of all
live contour nodes plus the N x r solve buffers, `qbhip.footprint`) fits in
the budget next to the windows already running; requests that do not fit
wait while smaller ones behind them may start (first fit).  Requests are
FP64 by default (on MI355X the faster path at these probe counts); a request
for the int8-residual mixed path larger than the whole budget runs with the
complex128-GEMM residual (24 instead of 56 bytes per matrix entry: no int8
residues of T), and a mixed one in FP64 (16 bytes), when that footprint fits;
one that fits in none is rejected with MemoryError.  The default budget is the device's free memory
at start-up less 12 GiB.  The host part of
each window (thin SVD, B, eig) runs on a thread pool, overlapping the next
windows' device work, with the BLAS threads shared out among the pool's
workers.  Device memory is allocated in stream order on each slot's stream,
so a slot's frees never wait for another slot's kernels.  Latency is measured from arrival to result.

With group=True, when a slot frees up, the pending requests that share the
admitted one's geometry and settings (shape, node counts, symmetry,
precision, nq, tolerances) join it, as many as fit in the budget (up to
`max_group`), and run as one window whose disks are assembled, factored and
solved as one batch (qb_window_create_multi) with one probe block; each
request still gets the eigenvalues of its own disk.  It is off by default:
from N ~ 1500 with 32 contour nodes one window's batch already fills the
device and the slots overlap windows, so grouping only coarsens the
scheduling (measured slower on the uneven-window workload).  Windows of
different N always run side by side in the slots.

A request with symmetry=True is served as the symmetry-reduced windows of its
shape's symmetry group (both reflections of the ellipse, stadium and disk,
the C_5 rotations of the star; _all_k_symm_CFIE_chebyshev!,
cheb-CFIE:698-824), and returns the union of the irrep spectra.  By default
(irreps_batched) the irreps are one window on a multi-irrep geometry
(qb_geom_set_orbits_multi): one assembly pass evaluates every image term for
all irreps, and one batched LU factors the G x nq reduced matrices.  With
irreps_batched=False every irrep is a job of its own, admitted by its own
footprint, and the request completes when its last irrep does.

No counterpart in the reference: its driver (compute_spectrum_beyn,
beyn.jl:778-921) loops over disks serially on the CPU.  The per-window
computation is the same construct_B_matrix + eigen!(B) as qbhip.beyn.
"""
from __future__ import annotations

import math
import os
import threading
import warnings
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

import dataclasses

from . import geometry as G
from .beyn import B_from_qr, Window, footprint, probe_block, probe_count
from .context import Context
from .solvers import CFIE_kress, Geometry, chebyshev_params
from .symmetry import Rotation, XYReflection, YReflection, node_divisor

SHAPES = {
    "circle": lambda: G.make_circle(1.0),
    "ellipse": lambda: G.make_full_ellipse(1.0, 0.6),
    "stadium": lambda: G.make_full_stadium(1.0, 1.0),
    "star": lambda: G.make_star_polar_boundary(1.0, 0.2, 5),
    "circle_hole": lambda: G.CircularHoleBilliard(1.0, holes=[(0.3, 0.35, 0.0)]),
}


# the irreps a symmetric request is split into, per shape
SYM_GROUPS = {
    "circle": lambda: [XYReflection(px, py) for px in (1, -1) for py in (1, -1)],
    "ellipse": lambda: [XYReflection(px, py) for px in (1, -1) for py in (1, -1)],
    "stadium": lambda: [XYReflection(px, py) for px in (1, -1) for py in (1, -1)],
    "star": lambda: [Rotation(5, m) for m in range(5)],
    "circle_hole": lambda: [YReflection(p) for p in (1, -1)],
}


# The Dirichlet eigenvalues are real.  Candidates of B whose imaginary part
# exceeds REAL_TOL |lam| come from directions of A0 above the rank threshold
# that carry no eigenvalue (spurious roots, the ones the reference's
# imag_k_check_EXPERIMENTAL, beyn.jl:623-721, residual-checks first); the
# count of eigenvalues reported by summarize() leaves them out.
REAL_TOL = 1e-8


# A request of tolerance tol solves T(z) X = V to a backward error of tol, so
# the singular values of A0 carry noise of about tol sigma_max(A0); rank
# directions below RANK_NOISE tol sigma_max are that noise (the reference's
# adaptive threshold, sigma_max 1e-15, beyn.jl:454, is the same rule for its
# FP64 solves).  With tol 1e-12 and sigma_max <= 1 this is the absolute
# svd_tol 1e-10 unchanged.
RANK_NOISE = 100.0

# A rank that still saturates with GROW_LIMIT times the starting probes (at
# least 64) is noise, not levels: the leak-margin Weyl count behind the
# starting probes already holds every level the filter passes.
GROW_LIMIT = 4


def rank_threshold(Rb: np.ndarray, svd_tol: float, tol: float, S=None) -> float:
    """Rank threshold of one moment block: max(svd_tol, RANK_NOISE tol
    sigma_max(A0)); sigma_max(A0) = sigma_max(R) for A0 = Q R."""
    if S is None:
        S = np.linalg.svd(Rb, compute_uv=False)
    smax = float(S[0]) if len(S) else 0.0
    return max(svd_tol, RANK_NOISE * tol * smax)


def real_roots(lam: np.ndarray) -> np.ndarray:
    """The candidates with |Im lam| <= REAL_TOL |lam|."""
    lam = np.asarray(lam)
    return lam[np.abs(lam.imag) <= REAL_TOL * np.abs(lam)]


@dataclass
class WindowRequest:
    rid: int
    shape: str
    k0: float
    R: float
    tol: float = 1e-12      # table tolerance and refinement backward error
    nq: int = 32
    ppw: float = 10.0       # points per wavelength: N = ppw k L / 2pi (rounded up, even)
    precision: str = "fp64"  # or "mixed" / "mixed_int8res" (complex64 LU + FP64 refinement)
    svd_tol: float = 1e-10
    seed: int = 0
    arrival: float = 0.0    # seconds after the stream starts
    N: Optional[List[int]] = None  # explicit node counts per component
    symmetry: bool = False  # serve as the symmetry-reduced windows of SYM_GROUPS[shape]
    r: Optional[int] = None  # starting probe count (default: the Weyl estimate, probe_count)


@dataclass
class WindowResult:
    rid: int
    lam: np.ndarray             # all eigenvalues of B
    inside: np.ndarray          # those with |lam - k0| <= R
    rk: int
    r: int
    N: int
    t_arrival: float
    t_admit: float
    t_done: float
    device_s: float
    host_s: float
    stats: dict = field(default_factory=dict)
    irreps: int = 1             # reduced windows merged into this result
    real: Optional[np.ndarray] = None  # inside ones with |Im lam| <= REAL_TOL |lam|

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


def node_counts(billiard: G.Billiard, k: float, ppw: float, mult: int = 2) -> List[int]:
    out = []
    for comp in billiard.full_boundary:
        L = sum(c.length for c in comp)
        n = int(math.ceil(ppw * k * L / (2 * math.pi)))
        out.append(max(16 + (-16) % mult, n + (-n) % mult))
    return out


class Server:
    def __init__(self, slots: int = 2, hbm_budget: Optional[int] = None, device: int = 0,
                 host_threads: int = 4, group: bool = False, max_group: int = 16,
                 irreps_batched: bool = True):
        self.slots = [Context(device) for _ in range(slots)]
        for c in self.slots:
            # windows of different sizes come and go: no per-slot buffer cache,
            # so the footprint accounting is the memory actually held
            c.set_cache_limit(0)
        self.free_slots = list(range(slots))
        self.group = bool(group)
        self.max_group = int(max_group)
        self.irreps_batched = bool(irreps_batched)
        if hbm_budget is None:
            free, _ = self.slots[0].mem_info()
            hbm_budget = max(free - 12 * 2 ** 30, 1 << 30)
        self.budget = int(hbm_budget)
        self.live_bytes = 0
        self.lock = threading.Condition()
        self.pending: List = []          # (req, future, t_arrival)
        self.geoms: Dict = {}            # (slot, shape, N tuple) -> Geometry
        self.pts_cache: Dict = {}        # (shape, N tuple) -> pts
        self.cheb_cache: Dict = {}       # (shape, N tuple, disks, nq, tol) -> cheb params
        self.prep_memo: Dict = {}        # id(job) -> (job, _prepare result)
        self.host_pool = ThreadPoolExecutor(max_workers=host_threads)
        # host stages of several windows run at once: give each its share of
        # the cores instead of letting every LAPACK call use all of them
        # (restored in close)
        self._blas_limit = None
        try:
            from threadpoolctl import threadpool_limits
            ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4
            self._blas_limit = threadpool_limits(limits=max(1, ncpu // host_threads),
                                                 user_api="blas")
        except Exception:
            self._blas_limit = None
        self.workers: List[threading.Thread] = []
        self.closed = False
        self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self.dispatcher.start()

    # -- public --------------------------------------------------------------
    def submit(self, req: WindowRequest, t_arrival: Optional[float] = None) -> Future:
        fut: Future = Future()
        ta = time.perf_counter() if t_arrival is None else t_arrival
        if not req.symmetry:
            jobs = [(req, None)]
        elif self.irreps_batched:
            # all irreps of the group in one window (qb_geom_set_orbits_multi)
            jobs = [(req, tuple(SYM_GROUPS[req.shape]()))]
        else:
            jobs = [(req, g) for g in SYM_GROUPS[req.shape]()]
        if req.symmetry and len(jobs) > 1:
            parts: List = [None] * len(jobs)
            left = [len(jobs)]
            guard = threading.Lock()

            def collect(i, f):
                with guard:
                    if fut.done():
                        return
                    if f.exception() is not None:
                        fut.set_exception(f.exception())
                        return
                    parts[i] = f.result()
                    left[0] -= 1
                    if left[0] == 0:
                        fut.set_result(_merge(parts))
        with self.lock:
            if self.closed:
                raise RuntimeError("server closed")
            for i, (r, g) in enumerate(jobs):
                f = fut if len(jobs) == 1 else Future()
                if len(jobs) > 1:
                    f.add_done_callback(lambda ff, i=i: collect(i, ff))
                self.pending.append(((r, g), f, ta))
            self.lock.notify_all()
        return fut

    def run_stream(self, requests: List[WindowRequest]) -> List:
        """Submit requests at their arrival offsets and wait for all.  Returns
        one entry per request in arrival order: its WindowResult, or the
        exception it ended with (e.g. MemoryError for a window too large)."""
        t0 = time.perf_counter()
        futs = []
        for req in sorted(requests, key=lambda r: r.arrival):
            dt = t0 + req.arrival - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            futs.append(self.submit(req, t0 + req.arrival))
        out = []
        for f in futs:
            try:
                out.append(f.result())
            except Exception as e:  # noqa: BLE001 - reported per request
                out.append(e)
        return out

    def warm_up(self, tune_N: int = 0, tune_r: int = 0, precisions=("fp64",)):
        """One small window (the ellipse (1, 0.6) at k0 = 60) in every slot at
        once, so the first request of each slot does not pay for loading the
        library's kernels into its device context.  With tune_N, each slot's
        context also times the GEMM tiles for the shape classes of a window of
        that size (qb_ctx_autotune with 32 nodes and tune_r probes, default
        N/10); the picks then serve windows of every size in those classes."""
        if tune_N:
            r = tune_r or max(64, tune_N // 10)
            for c in self.slots:
                for p in precisions:
                    c.autotune(int(tune_N), 32, int(r), p)
        futs = [self.submit(WindowRequest(-1 - i, "ellipse", 60.0, 0.3, tol=1e-8, nq=32))
                for i in range(len(self.slots))]
        for f in futs:
            f.result()

    def close(self):
        with self.lock:
            self.closed = True
            self.lock.notify_all()
        self.dispatcher.join()
        for w in self.workers:
            w.join()
        self.host_pool.shutdown(wait=True)
        if self._blas_limit is not None:
            self._blas_limit.restore_original_limits()
            self._blas_limit = None
        for g in self.geoms.values():
            g.close()
        for c in self.slots:
            c.close()

    # -- internals -----------------------------------------------------------
    def _prepare(self, job):
        memo = self.prep_memo.get(id(job))
        if memo is not None and memo[0] is job:
            return memo[1]
        prep = self._prepare_uncached(job)
        self.prep_memo[id(job)] = (job, prep)
        return prep

    def _prepare_uncached(self, job):
        req, sym = job
        irr = sym if isinstance(sym, tuple) else (sym,)
        sym = irr[0]
        b = SHAPES[req.shape]()
        mult = node_divisor(sym)
        Ns = (tuple(req.N) if req.N is not None
              else tuple(node_counts(b, req.k0 + req.R, req.ppw, mult)))
        key = (req.shape, Ns)
        if key not in self.pts_cache:
            self.pts_cache[key] = G.evaluate_points(CFIE_kress(req.ppw, b), b, req.k0, N=list(Ns))
        order = 1 if sym is None else sym.order
        # probes from the Weyl count of one symmetry sector
        bs = b if order == 1 else dataclasses.replace(b, area=b.area / order,
                                                      length=b.length / order)
        N = int(sum(Ns)) // order
        r = min(req.r if req.r is not None else probe_count(bs, req.k0, req.R, req.nq,
                                                             req.svd_tol), N)
        prec = req.precision
        nm = req.nq * len(irr)  # matrices of the window
        fp = footprint(N, nm, prec, r)
        if fp > self.budget and prec != "fp64":
            # the int8 residual keeps residues of T (32 more bytes per matrix
            # entry) and mixed precision keeps complex64 factors beside the
            # FP64 matrices: without them, when that is what fits
            for alt in (("mixed", "fp64") if prec == "mixed_int8res" else ("fp64",)):
                fa = footprint(N, nm, alt, r)
                if fa <= self.budget:
                    prec, fp = alt, fa
                    break
        return b, Ns, N, r, fp, prec

    def _dispatch(self):
        while True:
            with self.lock:
                while True:
                    if self.closed and not self.pending:
                        return
                    pick = None
                    if self.free_slots:
                        for idx, (job, fut, ta) in enumerate(self.pending):
                            prep = self._prepare(job)
                            fp = prep[4]
                            if fp > self.budget:
                                self.pending.pop(idx)
                                self.prep_memo.pop(id(job), None)
                                fut.set_exception(MemoryError(
                                    f"request {job[0].rid}: footprint {fp} B exceeds the budget {self.budget} B"))
                                pick = "rejected"
                                break
                            if self.live_bytes + fp <= self.budget:
                                pick = (idx, prep)
                                break
                    if pick == "rejected":
                        continue
                    if pick is not None:
                        idx, prep = pick
                        entries = [self.pending.pop(idx)]
                        if self.group:
                            prep = self._grow_group(entries, prep)
                        slot = self.free_slots.pop(0)
                        self.live_bytes += prep[4]
                        break
                    self.lock.wait(timeout=0.05)
            t = threading.Thread(target=self._run, args=(slot, entries, prep), daemon=True)
            self.workers.append(t)
            t.start()

    @staticmethod
    def _group_key(job, prep):
        req, sym = job
        return (req.shape, prep[1], sym, prep[5], req.nq, req.tol, req.svd_tol, req.ppw)

    def _grow_group(self, entries, prep):
        """Pending requests with the admitted one's geometry and settings join
        it while the batch fits in the budget (called with the lock held).
        Returns the group's prep: the largest probe count and the footprint
        of the batched window."""
        b, Ns, N, r, fp, prec = prep
        key = self._group_key(entries[0][0], prep)
        nq = entries[0][0][0].nq
        i = 0
        while i < len(self.pending) and len(entries) < self.max_group:
            job = self.pending[i][0]
            p2 = self._prepare(job)
            if self._group_key(job, p2) == key:
                r2 = max(r, p2[3])
                fp2 = footprint(N, nq * (len(entries) + 1), prec, r2)
                if self.live_bytes + fp2 <= self.budget and nq * (len(entries) + 1) <= 4096:
                    entries.append(self.pending.pop(i))
                    r, fp = r2, fp2
                    continue
            i += 1
        return b, Ns, N, r, fp, prec

    def _run(self, slot, entries, prep):
        """One window over the disks of `entries` (requests sharing a geometry)."""
        jobs = [e[0] for e in entries]
        with self.lock:
            for j in jobs:
                self.prep_memo.pop(id(j), None)
        req, sym = jobs[0]
        irr = sym if isinstance(sym, tuple) else (sym,)
        nirr = len(irr)
        b, Ns, N, r, fp, prec = prep
        reqs = [j[0] for j in jobs]
        t_admit = time.perf_counter()
        win = None
        try:
            ctx = self.slots[slot]
            pts = self.pts_cache[(req.shape, Ns)]
            solver = CFIE_kress(req.ppw, b, symmetry=irr[0])
            gkey = (slot, req.shape, Ns, sym)
            if gkey not in self.geoms:
                self.geoms[gkey] = Geometry(
                    pts, ctx, solver,
                    irreps=[CFIE_kress(req.ppw, b, symmetry=s_) for s_ in irr] if nirr > 1 else None)
            geom = self.geoms[gkey]
            th = np.exp(2j * np.pi * np.arange(req.nq) / req.nq)
            zj = np.concatenate([q.k0 + q.R * th for q in reqs])
            ck = (req.shape, Ns, tuple((round(q.k0, 3), round(q.R, 3)) for q in reqs), req.nq,
                  req.tol)
            tp = time.perf_counter()
            if ck not in self.cheb_cache:
                ttol = max(req.tol, 1e-13)
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore", RuntimeWarning)
                    nh, mh, nj, mj, errs = chebyshev_params(CFIE_kress(req.ppw, b), pts, zj,
                                                            tol=ttol, geom=geom)
                self.cheb_cache[ck] = (dict(n_panels_h=nh, M_h=mh, n_panels_j=nj, M_j=mj),
                                       bool(np.all(errs < ttol)))
            cheb, cheb_ok = self.cheb_cache[ck]
            t_cheb = time.perf_counter() - tp
            # the probe block is drawn on the host pool while the window is
            # assembled and factored (the C calls release the interpreter)
            vf = self.host_pool.submit(probe_block, np.random.default_rng(req.seed), N, r)
            tw = time.perf_counter()
            if len(reqs) == 1:
                win = Window(solver, pts, req.k0, req.R, req.nq, precision=prec, geom=geom,
                             **cheb)
            else:
                win = Window(solver, pts, [q.k0 for q in reqs], [q.R for q in reqs], req.nq,
                             precision=prec, geom=geom, **cheb)
            t_create = time.perf_counter() - tw
            tm = time.perf_counter()
            Rm, Cm = win.moments_qr(vf.result(), max(req.tol, 1e-13))
            t_moments = time.perf_counter() - tm
            if win.nblocks == 1:
                Rm, Cm = Rm[None], Cm[None]
            # Rank saturation: with rk == r every probe direction is used and
            # the disk may hold more eigenvalues than r probes resolve.  The
            # factors stay resident, so the moments are solved again with
            # twice the probes until the rank no longer saturates (the
            # reference retries with 2r, beyn.jl:291-309).  The extra solve
            # buffers are taken from the budget; when they do not fit, the
            # result is marked saturated instead.
            r_used, saturated, grow_rng = r, False, np.random.default_rng(req.seed + 7919)
            while _saturated(Rm, min(q.svd_tol for q in reqs), req.tol):
                r2 = min(2 * r_used, N)
                if r2 > GROW_LIMIT * max(r, 64):
                    saturated = True
                    break
                extra = footprint(N, req.nq * nirr * len(reqs), prec, r2) - fp
                if r2 == r_used or not self._reserve(extra):
                    saturated = True
                    break
                fp += extra
                Rm, Cm = win.moments_qr(probe_block(grow_rng, N, r2), max(req.tol, 1e-13))
                if win.nblocks == 1:
                    Rm, Cm = Rm[None], Cm[None]
                r_used = r2
            r = r_used
            st = dict(win.stats)
            st["precision_used"] = prec
            st["group"] = len(reqs)
            st["rank_saturated"] = saturated
            st["cheb_tol_met"] = cheb_ok
            st["phases_s"] = {"setup": round(tw - t_admit - t_cheb, 4), "cheb": round(t_cheb, 4),
                              "create": round(t_create, 4), "moments": round(t_moments, 4)}
        except Exception as e:  # release and report
            self._release(slot, fp)
            for _, fut, _ in entries:
                # no traceback: its frames would keep the window's buffers alive
                fut.set_exception(e.with_traceback(None))
            return
        finally:
            if win is not None:
                win.close()
        t_dev = time.perf_counter() - t_admit
        self._release(slot, fp)

        nwin = len(reqs)

        def host(c, q, t_arrival):
            # disk c: its block of every irrep (blocks are irrep-major)
            th = time.perf_counter()
            lams, rk = [], 0
            for e in range(nirr):
                B, Ur, S, rk_e = B_from_qr(Rm[e * nwin + c], Cm[e * nwin + c], q.svd_tol,
                                           RANK_NOISE * q.tol)
                lams.append(np.linalg.eigvals(B) if B.size else np.zeros(0, np.complex128))
                rk += rk_e
            lam = np.concatenate(lams)
            t_done = time.perf_counter()
            inside = np.sort_complex(lam[np.abs(lam - q.k0) <= q.R])
            st["symmetry"] = None if sym is None else repr(sym)
            return WindowResult(q.rid, lam, inside, rk, r * nirr, N * nirr, t_arrival, t_admit,
                                t_done, t_dev, t_done - th, st, irreps=nirr,
                                real=real_roots(inside))

        for c, ((q, _), fut, ta) in enumerate(entries):
            hf = self.host_pool.submit(host, c, q, ta)
            hf.add_done_callback(lambda f, fut=fut: fut.set_exception(f.exception())
                                 if f.exception() else fut.set_result(f.result()))

    def _reserve(self, extra: int) -> bool:
        """Take `extra` bytes of the budget for a running window if they fit."""
        with self.lock:
            if self.live_bytes + extra > self.budget:
                return False
            self.live_bytes += extra
            return True

    def _release(self, slot, fp):
        with self.lock:
            self.live_bytes -= fp
            self.free_slots.append(slot)
            self.free_slots.sort()
            self.lock.notify_all()


def _saturated(Rm: np.ndarray, svd_tol: float, tol: float) -> bool:
    """True when some moment block has full rank r at its threshold
    (rank_threshold): all r singular values of its R (A0 = Q R) clear it.
    sigma_min(R) <= min_i |R_ii| (R is triangular, its eigenvalues are its
    diagonal), so the SVD is only taken when every diagonal entry clears the
    absolute part of the threshold."""
    for Rb in Rm:
        if np.min(np.abs(np.diag(Rb))) < svd_tol:
            continue
        S = np.linalg.svd(Rb, compute_uv=False)
        if S[-1] >= rank_threshold(Rb, svd_tol, tol, S):
            return True
    return False


def _merge(parts: List[WindowResult]) -> WindowResult:
    """One result from the reduced windows of a symmetric request: the union of
    their spectra, completed when the last one completed."""
    p0 = parts[0]
    lam = np.concatenate([p.lam for p in parts])
    inside = np.sort_complex(np.concatenate([p.inside for p in parts]))
    return WindowResult(p0.rid, lam, inside, sum(p.rk for p in parts), sum(p.r for p in parts),
                        sum(p.N for p in parts), p0.t_arrival, min(p.t_admit for p in parts),
                        max(p.t_done for p in parts), sum(p.device_s for p in parts),
                        sum(p.host_s for p in parts), {"irreps": [p.stats for p in parts]},
                        irreps=len(parts), real=np.sort_complex(np.concatenate(
                            [p.real for p in parts])))


def summarize(results: List, t_span: Optional[float] = None) -> dict:
    """p50/p99 latency and eigenvalues/s over the completed windows (the real
    roots inside each disk; `candidates` counts all roots inside); entries
    that are exceptions (rejected requests) are counted separately."""
    failed = [r for r in results if not isinstance(r, WindowResult)]
    results = [r for r in results if isinstance(r, WindowResult)]
    lat = np.array([r.latency for r in results])
    n_eig = int(sum(len(r.real) for r in results))
    n_cand = int(sum(len(r.inside) for r in results))
    if t_span is None:
        t_span = max(r.t_done for r in results) - min(r.t_arrival for r in results)
    return {"windows": len(results), "rejected": len(failed), "eigenvalues": n_eig,
            "candidates": n_cand,
            "eigenvalues_per_s": n_eig / t_span if t_span > 0 else None,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "span_s": t_span}


def hole_node_counts(k: float, mult: int = 256) -> List[int]:
    """Node counts of the uneven-window workload on the unit disk with a hole of
    radius 0.3: N = 10 k L / 2 pi (L = 2.6 pi, the two circumferences) rounded
    up to a multiple of `mult`, split between the outer circle and the hole in
    the ratio of their lengths (10 : 3), both even."""
    n = int(math.ceil(10.0 * k * 2.6 * math.pi / (2 * math.pi) / mult)) * mult
    n_out = 2 * int(round(n * 10.0 / 13.0 / 2))
    return [n_out, n - n_out]


def uneven_windows(n: int, seed: int = 0, kmin: float = 100.0, kmax: float = 800.0,
                   R: float = 0.5, nq: int = 32, precision: str = "mixed",
                   tol: float = 1e-12) -> List[WindowRequest]:
    """The uneven batched-window workload: n windows on the unit disk with a
    hole of radius 0.3 centred at (0.35, 0), k0 seeded uniform in
    [kmin, kmax], N from hole_node_counts(k0 + R), all arriving at once."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k0 = float(rng.uniform(kmin, kmax))
        out.append(WindowRequest(rid=i, shape="circle_hole", k0=k0, R=R, tol=tol, nq=nq,
                                 precision=precision, seed=2000 + i, arrival=0.0,
                                 N=hole_node_counts(k0 + R)))
    return out


def poisson_stream(n: int, rate: float, seed: int = 0, shapes=("ellipse", "stadium", "star"),
                   kmin: float = 50.0, kmax: float = 1000.0, nq: int = 32,
                   precision: str = "fp64", symmetry: bool = False) -> List[WindowRequest]:
    """Seeded request stream of the serving workload: Poisson arrivals, shapes
    drawn from `shapes`, k0 log-uniform in [kmin, kmax], R uniform in
    [0.25, 0.5], tol 1e-8 or 1e-12."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(WindowRequest(rid=i, shape=str(rng.choice(shapes)),
                                 k0=float(np.exp(rng.uniform(np.log(kmin), np.log(kmax)))),
                                 R=float(rng.uniform(0.25, 0.5)),
                                 tol=float(rng.choice([1e-8, 1e-12])), nq=nq,
                                 precision=precision, seed=1000 + i, arrival=t,
                                 symmetry=symmetry))
    return out
