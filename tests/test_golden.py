"""The oracle and the host build of the device Bessel evaluator against the
committed golden fixtures (tests/golden/*.json, written by
tests/golden/make_golden.py from mpmath at 30 digits)."""
import ctypes as C
import json
import math
import os

import numpy as np
import scipy.special as sp

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HOSTB = os.path.join(ROOT, "tests", "native", "libqbt_bessel_host.so")


def _load(name):
    return json.load(open(os.path.join(GOLD, name)))


def _c(v):
    return complex(v[0], v[1])


def test_jm_zero_fixture_matches_oracle():
    g = _load("jm_zeros.json")
    ex = []
    for e in g["zeros"]:
        ex += [e["k"]] * e["mult"]
    np.testing.assert_allclose(O.jm_zeros_in_disk(g["center"], g["radius"]), sorted(ex),
                               rtol=1e-15)
    # SciPy's jn_zeros agrees with mpmath's besseljzero on every listed zero
    for e in g["interval_zeros"]:
        assert abs(sp.jn_zeros(e["m"], e["n"])[-1] - e["k"]) <= 1e-14 * e["k"]


def test_kress_R_fixture():
    g = _load("kress_R.json")
    for N, col in g.items():
        assert np.max(np.abs(O.kress_R_col(int(N)) - np.array(col))) < 1e-14


def test_table_samples_and_host_evaluator_match_mpmath_fixture():
    rows = _load("bessel_mpmath.json")
    z = np.array([_c(r["z"]) for r in rows])
    ref = {k: np.array([_c(r[k]) for r in rows]) for k in ("h0", "h1", "j0", "j1")}
    # the oracle's table samples (SciPy AMOS)
    for k, f in (("h0", lambda z: sp.hankel1(0, z)), ("h1", lambda z: sp.hankel1(1, z)),
                 ("j0", lambda z: sp.jv(0, z)), ("j1", lambda z: sp.jv(1, z))):
        assert np.max(np.abs(f(z) - ref[k]) / np.maximum(np.abs(ref[k]), 1e-3)) < 3e-14, k
    # the device evaluator, host build
    L = C.CDLL(HOSTB)
    zc = np.ascontiguousarray(z)
    outs = [np.empty_like(zc) for _ in range(4)]
    L.qbt_bessel_host(C.c_void_p(zc.ctypes.data), C.c_int64(len(zc)),
                      *[C.c_void_p(o.ctypes.data) for o in outs])
    j0, j1, h0, h1 = outs
    for k, got in (("h0", h0), ("h1", h1), ("j0", j0), ("j1", j1)):
        assert np.max(np.abs(got - ref[k]) / np.maximum(np.abs(ref[k]), 1e-3)) < 3e-14, k


def test_oracle_circle_spectrum_matches_multiplier_fixture():
    g = _load("circle_multipliers.json")
    N = 128
    comps = O.make_geometry("circle", [N])
    cp = O.ChebParams(6000, 7, 3000, 7)
    m_all = np.fft.fftfreq(N, 1.0 / N).astype(int)
    for key, ent in g.items():
        z = _c(ent["z"])
        T = O.cfie_assemble(comps, [z], cp)[0]
        F = O.dlp_assemble(comps[0], [z], cp)[0]
        lt, lf = np.fft.fft(T[:, 0]), np.fft.fft(F[:, 0])
        for e in ent["multipliers"]:
            i = int(np.flatnonzero(m_all == e["m"])[0])
            assert abs(lt[i] - _c(e["cfie"])) < 1e-10, (key, e["m"])
            assert abs(lf[i] - _c(e["dlp"])) < 1e-10, (key, e["m"])
