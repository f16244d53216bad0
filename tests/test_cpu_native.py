"""CPU tests of the native pieces: the C-ABI library exports what
include/qb.h declares, and the Bessel evaluator the device uses (host build)
matches AMOS (SciPy) and mpmath."""
import ctypes as C
import os
import re
import subprocess

import mpmath as mp
import numpy as np
import pytest
import scipy.special as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quantumbilliards.jl_amd", "libqbhip.so")
HOSTB = os.path.join(ROOT, "tests", "native", "libqbt_bessel_host.so")


def _declared():
    src = open(os.path.join(ROOT, "include", "qb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qb_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    import qbhip
    assert set(qbhip.EXPORTS) == set(_declared())
    L = C.CDLL(LIB)  # loads without a GPU
    for s in _declared():
        assert hasattr(L, s)


def _host_eval(z):
    L = C.CDLL(HOSTB)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    outs = [np.empty_like(z) for _ in range(4)]
    L.qbt_bessel_host(C.c_void_p(z.ctypes.data), C.c_int64(len(z)),
                      *[C.c_void_p(o.ctypes.data) for o in outs])
    return outs


@pytest.mark.parametrize("lo,hi", [(1e-6, 1e-3), (1e-3, 0.2), (0.2, 2), (2, 20), (20, 200),
                                   (200, 2600)])
def test_host_bessel_matches_amos(lo, hi):
    rng = np.random.default_rng(int(lo * 1e6) + 7)
    rho = np.exp(rng.uniform(np.log(lo), np.log(hi), 20000))
    im = rng.uniform(-1.5, 1.5, rho.size) * np.minimum(1.0, rho)
    z = np.sqrt(np.maximum(rho ** 2 - im ** 2, 1e-300)) + 1j * im
    j0, j1, h0, h1 = _host_eval(z)
    for a, b in ((h0, sp.hankel1(0, z)), (h1, sp.hankel1(1, z))):
        assert np.max(np.abs(a - b) / np.abs(b)) < 2e-14
    scale = np.minimum(1.0, np.exp(np.abs(z.imag)) / np.sqrt(np.maximum(np.abs(z), 1.0)))
    for a, b in ((j0, sp.jv(0, z)), (j1, sp.jv(1, z))):
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), scale)) < 1e-14


def test_host_bessel_matches_mpmath_spot_values():
    zs = np.array([1e-4 + 1e-5j, 0.15 - 0.01j, 1.7 + 0.3j, 7.77 - 0.4j, 19.9 + 0.2j, 20.1 - 0.2j,
                   400.5 + 0.45j, 2047.3 - 1.1j])
    j0, j1, h0, h1 = _host_eval(zs)
    mp.mp.dps = 30
    for i, z in enumerate(zs):
        w = mp.mpc(z)
        for got, ref in ((h0[i], mp.hankel1(0, w)), (h1[i], mp.hankel1(1, w)),
                         (j0[i], mp.besselj(0, w)), (j1[i], mp.besselj(1, w))):
            ref = complex(ref)
            assert abs(got - ref) <= 2e-14 * max(abs(ref), 1e-3), (z, got, ref)


def test_host_bessel_region_boundaries_are_continuous():
    # values just below and above the series / Miller / asymptotic switches
    for rho in (2.0, 20.0):
        z = np.array([rho * (1 - 4e-16) + 0.1j, rho * (1 + 4e-16) + 0.1j])
        j0, j1, h0, h1 = _host_eval(z)
        for v in (j0, j1, h0, h1):
            assert abs(v[0] - v[1]) < 2e-14 * max(abs(v[0]), 0.2)
