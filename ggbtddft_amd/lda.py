"""Spin-unpolarised LDA on the host (used by the ground-state precompute).

Exchange: Slater v_x = -(3 rho / pi)^{1/3} (dft.py:9-20 of the reference).
Correlation: Perdew-Wang 1992 (Phys. Rev. B 45, 13244), parameters
A = 0.031091, alpha1 = 0.21370, beta1..4 = 7.5957, 3.5876, 1.6382, 0.49294.
The device kernel ``lda_vxc`` in csrc/ggbtd.hip evaluates the same formula.
"""
import numpy as np

RHO_MIN = 1e-40
_A, _A1 = 0.031091, 0.21370
_B = (7.5957, 3.5876, 1.6382, 0.49294)


def vxc(rho):
    rho = np.asarray(rho, dtype=np.float64)
    out = np.zeros_like(rho)
    sel = rho > RHO_MIN
    n = rho[sel]
    rs = np.cbrt(3.0 / (4.0 * np.pi * n))
    s = np.sqrt(rs)
    den = 2.0 * _A * (_B[0] * s + _B[1] * rs + _B[2] * rs * s + _B[3] * rs * rs)
    dden = _A * (_B[0] / s + 2.0 * _B[1] + 3.0 * _B[2] * s + 4.0 * _B[3] * rs)
    pre = -2.0 * _A * (1.0 + _A1 * rs)
    lg = np.log1p(1.0 / den)
    ec = pre * lg
    dec = -2.0 * _A * _A1 * lg - pre * dden / (den * den + den)
    out[sel] = -np.cbrt(3.0 * n / np.pi) + ec - rs * dec / 3.0
    return out
