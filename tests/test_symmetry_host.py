"""Symmetry reduction on the CPU: the oracle's reduced assembly pinned by what
the mathematics fixes, and the product's host-side orbit tables checked
against the oracle's independent construction.

Pins of the oracle (oracle.reduce_matrix, orbits_index, orbits_coords):
  * invariance: for a geometry invariant under the group, the full operator
    commutes with the node permutations, so the symmetry-adapted columns S
    (S[img[b, l], b] = scale[b, l]) span an invariant subspace and
    A S = S A_red holds to rounding -- a wrong weight, a transposed sum or a
    wrong image breaks it;
  * the spectra of the reduced matrices over all irreps together are the
    spectrum of the full matrix;
  * on the unit disk the Beyn eigenvalues of irrep m of C_4 are the zeros of
    J_|m'| over the integers m' = m (mod 4), and those of the y -> -y
    reflection with parity +1 / -1 the zeros of J_m' for m' >= 0 / m' >= 1
    (cos / sin modes), each once (tests/golden/jm_zeros.json).
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CP = O.ChebParams(3000, 7, 1500, 7)


def _S(img, scale, N):
    m, G = img.shape
    S = np.zeros((N, m), np.complex128)
    for b in range(m):
        S[img[b], b] = scale[b]
    return S


def _cases():
    # (name, comps, op, kind, n, m, parity)
    ell = O.make_geometry("ellipse", [120])
    star = O.make_geometry("star", [200])
    hole = O.make_geometry("circle_hole", [96, 32], hole=(0.3, 0.0, 0.0))
    return [
        ("cfie-ellipse-origin", ell, "cfie", "origin", 0, 0, (-1, 1)),
        ("cfie-ellipse-yaxis", ell, "cfie", "y_axis", 0, 0, -1),
        ("cfie-star-rot5", star, "cfie", "rot", 5, 2, 1),
        ("cfie-hole-rot4", hole, "cfie", "rot", 4, 1, 1),
        ("cfie-hole-xaxis", hole, "cfie", "x_axis", 0, 0, 1),
        ("dlp-ellipse-origin", ell, "dlp", "origin", 0, 0, (1, -1)),
        ("dlp-star-rot5", star, "dlp", "rot", 5, 3, 1),
    ]


def _full(comps, op, z):
    if op == "cfie":
        return O.cfie_assemble(comps, [z], CP)[0]
    return O.dlp_assemble(comps[0], [z], CP)[0]


def _orbits(comps, op, kind, n, m, p):
    N = sum(len(c.x) for c in comps)
    if op == "dlp":
        return O.orbits_index(N, kind, n, m, p)
    return O.orbits_coords(comps, kind, n, m, p)


@pytest.mark.parametrize("case", range(7))
def test_oracle_reduction_is_an_invariant_subspace(case):
    name, comps, op, kind, n, m, p = _cases()[case]
    z = 14.0 + 0.2j
    A = _full(comps, op, z)
    img, sc = _orbits(comps, op, kind, n, m, p)
    N = A.shape[0]
    assert sorted(img.ravel().tolist()) == list(range(N)), name  # a partition of the nodes
    Ar = O.reduce_matrix(A, img, sc)
    S = _S(img, sc, N)
    err = np.max(np.abs(A @ S - S @ Ar)) / np.max(np.abs(A))
    assert err < 1e-12, (name, err)
    # a plausible mistake -- conjugated weights -- is caught
    if kind == "rot" and m not in (0, n // 2 if n % 2 == 0 else -1):
        bad = O.reduce_matrix(A, img, sc.conj())
        assert np.max(np.abs(A @ S - S @ bad)) / np.max(np.abs(A)) > 1e-6


def test_oracle_irreps_together_carry_the_full_spectrum():
    comps = O.make_geometry("circle_hole", [96, 32], hole=(0.3, 0.0, 0.0))
    z = 9.0 - 0.1j
    A = O.cfie_assemble(comps, [z], CP)[0]
    ev = []
    for m in range(4):
        img, sc = O.orbits_coords(comps, "rot", 4, m)
        ev.append(np.linalg.eigvals(O.reduce_matrix(A, img, sc)))
    got = np.sort_complex(np.concatenate(ev))
    ref = np.sort_complex(np.linalg.eigvals(A))
    # match as multisets (sorting can pair near-equal values differently)
    used = np.zeros(len(ref), bool)
    worst = 0.0
    for g in got:
        d = np.abs(ref - g) + used * 1e300
        i = int(np.argmin(d))
        used[i] = True
        worst = max(worst, d[i])
    assert worst < 1e-10


def _irrep_zeros(rule):
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "jm_zeros.json")))
    out = []
    for e in g["zeros"]:
        for mm in ({e["m"], -e["m"]} if e["m"] else {0}):
            if rule(mm):
                out.append(e["k"])
    return np.sort(out)


def _beyn_reduced(comps, op, kind, n, m, p, r=12):
    img, sc = _orbits(comps, op, kind, n, m, p)
    rng = np.random.default_rng(3)
    V = (rng.standard_normal((len(img), r)) + 1j * rng.standard_normal((len(img), r))) * np.sqrt(0.5)
    asm = lambda zs: np.stack([O.reduce_matrix(T, img, sc) for T in O.cfie_assemble(comps, zs)])
    lam, S, rk, _, _ = O.beyn_solve(asm, 20.0, 0.5, 32, V, 1e-10)
    return np.sort(lam[np.abs(lam - 20) <= 0.5].real)


@pytest.mark.parametrize("m", range(4))
def test_oracle_reduced_beyn_on_the_disk_c4(m):
    comps = O.make_geometry("circle", [256])
    got = _beyn_reduced(comps, "cfie", "rot", 4, m, 1)
    ex = _irrep_zeros(lambda mm: (mm - m) % 4 == 0)
    assert len(got) == len(ex) and np.max(np.abs(got - ex) / ex) < 1e-12, (got, ex)


@pytest.mark.parametrize("p", [1, -1])
def test_oracle_reduced_beyn_on_the_disk_reflection(p):
    comps = O.make_geometry("circle", [256])
    got = _beyn_reduced(comps, "cfie", "x_axis", 0, 0, p)
    ex = _irrep_zeros(lambda mm: mm >= (0 if p == 1 else 1))
    assert len(got) == len(ex) and np.max(np.abs(got - ex) / ex) < 1e-12, (got, ex)


# ---------------------------------------------------------------------------
# the product's host logic (no device)
# ---------------------------------------------------------------------------
def _q():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "quantumbilliards.jl_amd"))
    import qbhip
    return qbhip


def _pts_from(comps):
    Q = _q()
    out = []
    for c in comps:
        out.append(Q.BoundaryPointsCFIE(np.column_stack([c.x, c.y]), np.column_stack([c.dx, c.dy]),
                                        np.column_stack([c.ddx, c.ddy]), c.ts, c.ws))
    return out


def _sym(Q, kind, n, m, p):
    if kind == "rot":
        return Q.Rotation(n, m)
    if kind == "origin":
        return Q.XYReflection(*p)
    return Q.XReflection(p) if kind == "y_axis" else Q.YReflection(p)


@pytest.mark.parametrize("case", range(7))
def test_product_orbit_tables_match_the_oracle(case):
    Q = _q()
    name, comps, op, kind, n, m, p = _cases()[case]
    img, sc = _orbits(comps, op, kind, n, m, p)
    pts = _pts_from(comps)
    sym = _sym(Q, kind, n, m, p)
    if op == "dlp":
        pimg, psc = Q.dlp_orbits(len(pts[0]), sym)
    else:
        pimg, psc = Q.cfie_orbits(pts, sym)
    assert np.array_equal(pimg, img), name
    assert np.max(np.abs(psc - sc)) < 1e-15, name


def test_product_node_rule_and_matrix_size():
    Q = _q()
    b = Q.make_star_polar_boundary(1.0, 0.2, 5)
    for sym, div in ((None, 2), (Q.Rotation(5, 0), 10), (Q.Rotation(4, 1), 4),
                     (Q.XReflection(1), 4), (Q.XYReflection(1, -1), 4)):
        s = Q.CFIE_kress(8.0, b, symmetry=sym)
        for k in (11.0, 23.7, 40.2):
            pts = Q.evaluate_points(s, b, k)
            N = len(pts[0])
            raw = max(20, int(round(k * b.length * 8.0 / (2 * np.pi))))
            assert N % div == 0 and raw <= N < raw + div
            order = 1 if sym is None else sym.order
            assert Q.boundary_matrix_size_for_solver(s, pts) == N // order
    # DLP: XReflection and YReflection share the index map q -> N - q + 1
    a, _ = Q.dlp_orbits(16, Q.XReflection(1))
    c, _ = Q.dlp_orbits(16, Q.YReflection(1))
    assert np.array_equal(a, c) and a[0].tolist() == [0, 15]
    # N = 122 puts node 31 at t = pi/2, on the y axis: the reflection maps it
    # onto itself and its orbit repeats a node (the library rejects such a
    # table, see tests/test_gpu_symmetry.py)
    ell = O.make_geometry("ellipse", [122])
    img, _ = O.orbits_coords(ell, "origin", parity=(1, 1))
    assert any(len(set(r)) < len(r) for r in img.tolist())
    pimg, _ = Q.cfie_orbits(_pts_from(ell), Q.XYReflection(1, 1))
    assert np.array_equal(pimg, img)
