"""LDA exchange-correlation potential (oracle; test infrastructure only).

Exchange: Slater, v_x = -(3 rho/pi)^{1/3}; the same expression the reference
uses (``dft.py:9-20``, ``h_calculateExchangeCorrelationPotential.cuh:10-11``).

Correlation: Perdew & Wang, Phys. Rev. B 45, 13244 (1992), spin-unpolarised
G(r_s; A, a1, b1..b4) with p = 1 and the published constants
A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294.
The reference uses Chachiyo correlation instead (``dft.py:58-87``); PW92 is the
choice of this project.

  eps_c = -2A(1 + a1 rs) ln(1 + 1/(2A(b1 rs^1/2 + b2 rs + b3 rs^3/2 + b4 rs^2)))
  v_c   = eps_c - rs/3 d eps_c/d rs

Densities at or below RHO_MIN give v_xc = 0 (both the oracle and the kernels
apply the same cut; at 1e-40 the dropped exchange is ~2e-14 Hartree).

Pinned by: v = d(rho eps)/d rho (numerical derivative, catches a wrong
derivative term); the uniform-gas correlation energies of the independent
Perdew-Zunger 1981 fit of the same Ceperley-Alder data (agreement to a few
1e-4 Hartree); the Slater closed form.
"""
import numpy as np

RHO_MIN = 1e-40

PW92_A = 0.031091
PW92_A1 = 0.21370
PW92_B1 = 7.5957
PW92_B2 = 3.5876
PW92_B3 = 1.6382
PW92_B4 = 0.49294


def rs_of_rho(rho):
    return np.cbrt(3.0 / (4.0 * np.pi * rho))


def eps_x(rho):
    return -0.75 * np.cbrt(3.0 * rho / np.pi)


def v_x(rho):
    return -np.cbrt(3.0 * rho / np.pi)


def eps_c_pw92(rho):
    rs = rs_of_rho(rho)
    srs = np.sqrt(rs)
    q1 = 2.0 * PW92_A * (PW92_B1 * srs + PW92_B2 * rs + PW92_B3 * rs * srs + PW92_B4 * rs * rs)
    return -2.0 * PW92_A * (1.0 + PW92_A1 * rs) * np.log1p(1.0 / q1)


def v_c_pw92(rho):
    rs = rs_of_rho(rho)
    srs = np.sqrt(rs)
    q0 = -2.0 * PW92_A * (1.0 + PW92_A1 * rs)
    q1 = 2.0 * PW92_A * (PW92_B1 * srs + PW92_B2 * rs + PW92_B3 * rs * srs + PW92_B4 * rs * rs)
    dq1 = PW92_A * (PW92_B1 / srs + 2.0 * PW92_B2 + 3.0 * PW92_B3 * srs + 4.0 * PW92_B4 * rs)
    lg = np.log1p(1.0 / q1)
    eps = q0 * lg
    deps = -2.0 * PW92_A * PW92_A1 * lg - q0 * dq1 / (q1 * q1 + q1)
    return eps - rs / 3.0 * deps


def v_xc(rho):
    """Slater + PW92 potential, zero where rho <= RHO_MIN."""
    rho = np.asarray(rho, dtype=np.float64)
    out = np.zeros_like(rho)
    ok = rho > RHO_MIN
    rr = rho[ok]
    out[ok] = v_x(rr) + v_c_pw92(rr)
    return out


def eps_xc(rho):
    rho = np.asarray(rho, dtype=np.float64)
    out = np.zeros_like(rho)
    ok = rho > RHO_MIN
    rr = rho[ok]
    out[ok] = eps_x(rr) + eps_c_pw92(rr)
    return out
