"""Host logic of the spectrum driver (no GPU): Weyl window planning and the
oracle's residual check, pinned by what the mathematics fixes."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from qbhip import geometry as G
from qbhip.spectrum import beyn_disks_from_windows, delta_weyl, plan_weyl_windows


@pytest.mark.parametrize("k1,k2,m,Rmax", [(100.0, 130.0, 50, 0.5), (10.0, 40.0, 10, 1.0),
                                          (500.0, 500.7, 300, 0.5), (20.0, 20.3, 10, 0.5)])
def test_weyl_windows_cover_interval(k1, k2, m, Rmax):
    """Windows tile [k1, k2] without gaps, none longer than 2 Rmax, and each
    window not cut by the length cap or the interval end holds m levels of
    the area term of Weyl's law, A((k+dk)^2 - k^2)/4pi, to 0.1 level."""
    b = G.make_full_ellipse(1.0, 0.6)
    iv = plan_weyl_windows(b, k1, k2, m=m, Rmax=Rmax, fundamental=False)
    assert iv[0][0] == k1 and abs(iv[-1][1] - k2) < 1e-9
    for (a0, b0), (a1, _) in zip(iv, iv[1:]):
        assert a1 == b0
    for a, c in iv:
        assert c > a and c - a <= 2 * Rmax * (1 + 1e-12)
        n = b.area * (c * c - a * a) / (4 * math.pi)
        assert abs(n - delta_weyl(b, a, c - a, False)) < 1e-9 * max(n, 1.0)
        if c - a < 2 * Rmax * (1 - 1e-9) and c < k2 - 1e-9:
            assert abs(n - m) <= 0.1 + 1e-9
    k0, R = beyn_disks_from_windows(iv)
    assert np.allclose(k0.real - R, [a for a, _ in iv]) and np.allclose(k0.real + R,
                                                                         [c for _, c in iv])
    assert np.all(k0.imag == 0)


def test_weyl_windows_short_tail_is_one_window():
    """An interval holding fewer than m levels and shorter than 2 Rmax is a
    single window."""
    b = G.make_circle(1.0)
    assert plan_weyl_windows(b, 20.0, 20.4, m=50, Rmax=0.5) == [(20.0, 20.4)]


def test_oracle_residual_select_on_circle_modes():
    """On the circle T(z) is circulant, so the Fourier mode e^{i m t} is an
    eigenvector with eigenvalue mu_m(z); at z = j_{m,n} (a Dirichlet
    eigenvalue) mu_m vanishes up to the discretisation error, away from it it
    does not.  The oracle's residual check must keep the first and drop the
    second, and its operator 1-norm must equal NumPy's."""
    from scipy.special import jn_zeros
    N = 128
    comps = O.make_geometry("circle", [N])
    cp = O.ChebParams(2000, 5, 800, 5)
    t = comps[0].ts
    z_eig = jn_zeros(6, 4)[-1]          # j_{6,4} = 20.3208
    z_off = z_eig + 0.05
    lam = np.array([z_eig, z_off], dtype=complex)
    Uk = np.column_stack([np.exp(6j * t) / math.sqrt(N)] * 2)
    Y = np.eye(2, dtype=complex)
    T_of = lambda z: O.cfie_assemble(comps, [z], cp)[0]
    idx, tens, tensN, det = O.residual_select(T_of, lam, Uk, Y, 20.3 + 0j, 0.5, res_tol=1e-8)
    assert list(idx) == [0]
    assert tens[0] < 1e-10 and det[1][0] > 1e-3
    T = T_of(z_eig)
    assert abs(det[0][3] - np.linalg.norm(T, 1)) <= 1e-13 * np.linalg.norm(T, 1)
    assert abs(det[0][4] - np.linalg.norm(T, np.inf)) <= 1e-13 * np.linalg.norm(T, np.inf)
