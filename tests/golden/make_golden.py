"""Writes the golden fixtures of this directory from closed forms evaluated in
mpmath at 30 significant digits (independent of SciPy/AMOS and of this
build's code).  Run from the repository root:  python tests/golden/make_golden.py

  jm_zeros.json            Dirichlet eigenvalues of the unit disk (zeros of J_m,
                           multiplicity 2 for m > 0) in |k - 20| <= 0.5 and in
                           [19.2, 20.8]
  kress_R.json             Kress log-quadrature weights R_d (d = i - j mod N),
                           brute-force cosine sums for N = 8, 9, 16, 64
  bessel_mpmath.json       H0^(1), H1^(1), J0, J1 at complex points covering the
                           series, recurrence and asymptotic regions
  circle_multipliers.json  Fourier multipliers of the unit-circle CFIE and DLP
                           operators, pi z J_m (H_m - i H_m') and -i pi z J_m H_m',
                           m = -20..20
  circle_r07_multipliers.json
                           the same on the circle of radius a = 0.7 (speed a,
                           curvature 1/a): pi a z J_m(a z) (H_m(a z) - i H_m'(a z))
                           and -i pi a z J_m(a z) H_m'(a z)
  annulus_zeros.json       Dirichlet eigenvalues of the concentric annulus
                           0.3 < |x| < 1 in |k - 14| <= 0.5: zeros of
                           J_m(k) Y_m(0.3 k) - J_m(0.3 k) Y_m(k), multiplicity 2
                           for m > 0
"""
import json
import os

import mpmath as mp

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))


def c2(z):
    z = mp.mpc(z)
    return [float(z.real), float(z.imag)]


def jm_zeros(lo, hi):
    out = []
    for m in range(0, 80):
        n = 1
        while True:
            z = mp.besseljzero(m, n)
            if z > hi:
                break
            if z >= lo:
                out.append({"m": m, "n": n, "k": float(z), "mult": 1 if m == 0 else 2})
            n += 1
        if mp.besseljzero(m, 1) > hi:
            break
    return sorted(out, key=lambda e: e["k"])


def kress_R(N):
    n = N // 2
    mmax = n - 1 if N % 2 == 0 else (N - 1) // 2
    col = []
    for d in range(N):
        s = mp.mpf(0)
        for m in range(1, mmax + 1):
            s += mp.cos(2 * mp.pi * m * d / N) / m
        v = -(4 * mp.pi / N) * s
        if N % 2 == 0:
            v -= (4 * mp.pi / N ** 2) * (-1) ** d
        col.append(float(v))
    return col


def bessel_points():
    pts = []
    for rho in (1e-3, 0.05, 0.3, 1.0, 1.9, 2.1, 5.0, 12.0, 19.5, 20.5, 60.0, 400.0, 2500.0):
        for im in (-1.2, -0.3, 0.0, 0.4, 1.5):
            if abs(im) >= rho:
                continue
            re = float(mp.sqrt(rho * rho - im * im))
            pts.append(mp.mpc(re, im))
    return pts


def circle_multipliers(a, zs):
    out = {}
    for z in zs:
        ent = []
        w = a * z
        for m in range(-20, 21):
            J = mp.besselj(m, w)
            H = mp.hankel1(m, w)
            Hp = (mp.hankel1(m - 1, w) - mp.hankel1(m + 1, w)) / 2
            ent.append({"m": m, "cfie": c2(mp.pi * w * J * (H - 1j * Hp)),
                        "dlp": c2(-1j * mp.pi * w * J * Hp)})
        out[f"{float(z.real)}{float(z.imag):+}j"] = {"z": c2(z), "radius": float(a),
                                                      "multipliers": ent}
    return out


def annulus_zeros(rho, lo, hi, mmax=60):
    """Zeros in [lo, hi] of f_m(k) = J_m(k) Y_m(rho k) - J_m(rho k) Y_m(k):
    sign changes on a fine grid, then mpmath's root finder."""
    f = lambda m, k: (mp.besselj(m, k) * mp.bessely(m, rho * k)
                      - mp.besselj(m, rho * k) * mp.bessely(m, k))
    out = []
    for m in range(mmax):
        ks = [lo - 0.05 + i * (hi - lo + 0.1) / 2000 for i in range(2001)]
        vs = [f(m, mp.mpf(k)) for k in ks]
        for i in range(len(ks) - 1):
            if vs[i] * vs[i + 1] < 0:
                z = mp.findroot(lambda k: f(m, k), (mp.mpf(ks[i]), mp.mpf(ks[i + 1])),
                                solver="anderson")
                if lo <= z <= hi:
                    out.append({"m": m, "k": float(z), "mult": 1 if m == 0 else 2})
    return sorted(out, key=lambda e: e["k"])


def main():
    disk = jm_zeros(19.5, 20.5)
    disk = [e for e in disk if abs(e["k"] - 20.0) <= 0.5]
    json.dump({"center": 20.0, "radius": 0.5, "zeros": disk,
               "interval": [19.2, 20.8], "interval_zeros": jm_zeros(19.2, 20.8)},
              open(os.path.join(HERE, "jm_zeros.json"), "w"), indent=1)
    json.dump({str(N): kress_R(N) for N in (8, 9, 16, 64)},
              open(os.path.join(HERE, "kress_R.json"), "w"), indent=1)
    rows = []
    for z in bessel_points():
        rows.append({"z": c2(z), "h0": c2(mp.hankel1(0, z)), "h1": c2(mp.hankel1(1, z)),
                     "j0": c2(mp.besselj(0, z)), "j1": c2(mp.besselj(1, z))})
    json.dump(rows, open(os.path.join(HERE, "bessel_mpmath.json"), "w"), indent=1)
    mult = {}
    for z in (mp.mpc(9.0, 0.2), mp.mpc(11.0, -0.3)):
        ent = []
        for m in range(-20, 21):
            J = mp.besselj(m, z)
            H = mp.hankel1(m, z)
            Hp = (mp.hankel1(m - 1, z) - mp.hankel1(m + 1, z)) / 2
            ent.append({"m": m, "cfie": c2(mp.pi * z * J * (H - 1j * Hp)),
                        "dlp": c2(-1j * mp.pi * z * J * Hp)})
        mult[f"{float(z.real)}{float(z.imag):+}j"] = {"z": c2(z), "multipliers": ent}
    json.dump(mult, open(os.path.join(HERE, "circle_multipliers.json"), "w"), indent=1)
    json.dump(circle_multipliers(mp.mpf("0.7"), (mp.mpc(9.0, 0.2), mp.mpc(13.0, -0.25))),
              open(os.path.join(HERE, "circle_r07_multipliers.json"), "w"), indent=1)
    json.dump({"inner_radius": 0.3, "center": 14.0, "radius": 0.5,
               "zeros": annulus_zeros(mp.mpf("0.3"), 13.5, 14.5)},
              open(os.path.join(HERE, "annulus_zeros.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
