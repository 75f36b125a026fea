"""Pins the oracle's Hartree multipoles (oracle/hartree.py and the potential
pipeline of oracle/propagate.py) against closed forms and the reference's
direct-sum formula."""
import numpy as np
from scipy.special import exp1

from oracle import angular as A, hartree as H, mesh as M


def test_scan_equals_direct_sum():
    rng = np.random.default_rng(1)
    m = M.lobatto_mesh(150, 40.0, 1.0)
    q = rng.normal(size=(25, 150)) * np.exp(-m["r"] / 4.0)
    Vd = H.direct(q, m["r"])
    Vs = H.scan(q, m["r"])
    assert np.abs(Vd - Vs).max() < 1e-13 * np.abs(Vd).max()


def test_hydrogen_1s_closed_form_second_order():
    errs = []
    for N in (128, 256):
        m = M.lobatto_mesh(N, 40.0, 1.0)
        r = m["r"]
        q0 = 4.0 * r * r * np.exp(-2.0 * r) * m["wr"]
        V = H.scan(q0[None], r)[0]
        errs.append(np.abs(V - (1.0 / r - (1.0 + 1.0 / r) * np.exp(-2.0 * r))).max())
    assert errs[1] < 2e-4
    assert 3.0 < errs[0] / errs[1] < 5.0     # O(N^-2) kink quadrature


def test_quadrupole_shell_closed_form():
    # rho(r, theta) = e^{-r} P_2(cos theta):
    # V(r, theta) = 4 pi/5 P_2 [r^-3 int_0^r e^-s s^4 ds + r^2 E_1(r)]
    N, lmax = 384, 4
    m = M.lobatto_mesh(N, 60.0, 1.0)
    r = m["r"]
    chi, om = A.theta_grid(2 * lmax + 1)
    PL = A.legendre_table(2 * lmax, chi)
    rho = np.exp(-r)[:, None] * PL[None, :, 2]
    rho_u = rho * (2.0 * np.pi * om[None, :] * (m["wr"] * r * r)[:, None])
    q = (rho_u @ PL).T
    V = H.scan(q, r)
    inner = 24.0 - np.exp(-r) * (r ** 4 + 4 * r ** 3 + 12 * r ** 2 + 24 * r + 24)
    exact = 4.0 * np.pi / 5.0 * (inner / r ** 3 + r * r * exp1(r))
    sel = r < 20.0
    assert np.abs(V[2][sel] - exact[sel]).max() < 2e-3 * np.abs(exact).max()
    assert np.abs(np.delete(V, 2, axis=0)).max() < 1e-10
