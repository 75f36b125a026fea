"""Pins the oracle's LDA (oracle/xc.py): potentials are derivatives of the
energy density, correlation agrees with the independent Perdew-Zunger 1981
fit of the same Ceperley-Alder data, exchange is Slater's closed form."""
import numpy as np
import pytest

from oracle import xc


def _pz81(rs):
    rs = np.asarray(rs, dtype=float)
    hi = -0.1423 / (1.0 + 1.0529 * np.sqrt(rs) + 0.3334 * rs)
    lo = 0.0311 * np.log(rs) - 0.048 + 0.0020 * rs * np.log(rs) - 0.0116 * rs
    return np.where(rs >= 1.0, hi, lo)


@pytest.mark.parametrize("rs", [0.5, 1.0, 2.0, 5.0, 10.0, 20.0])
def test_pw92_vs_pz81(rs):
    rho = 3.0 / (4.0 * np.pi * rs ** 3)
    assert abs(xc.eps_c_pw92(rho) - _pz81(rs)) < 6e-4


@pytest.mark.parametrize("rho", [1e-6, 1e-3, 0.1, 1.0, 30.0, 1e4])
def test_potentials_are_energy_derivatives(rho):
    h = 1e-5 * rho
    f = lambda n: n * (xc.eps_x(n) + xc.eps_c_pw92(n))
    num = (f(rho + h) - f(rho - h)) / (2 * h)
    assert abs(num - (xc.v_x(rho) + xc.v_c_pw92(rho))) < 1e-7 * max(1.0, abs(num))


def test_slater_closed_form_and_cut():
    rho = np.array([0.0, 1e-41, 0.2, 8.0])
    v = xc.v_xc(rho)
    assert v[0] == 0.0 and v[1] == 0.0
    assert abs(xc.v_x(8.0) + (24.0 / np.pi) ** (1.0 / 3.0)) < 1e-14
    assert np.all(v[2:] < 0.0)
