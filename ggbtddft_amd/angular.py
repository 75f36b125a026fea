"""Angular tables (precompute, host FP64).

Orbital k with magnetic number m >= 0 is expanded in Y_lm, l = m..Lmax
(m conserved under linear polarisation).  The theta grid is Gauss-Legendre in
x = cos(theta); the l -> theta transform of |m| is
    B_m[j, l] = sqrt(omega_j) Theta_lm(x_j)
with Theta_lm the L2([-1,1])-normalised associated Legendre functions, built by
the recurrence of the cos(theta) matrix elements
    c(l, m) = sqrt(((l+1)^2 - m^2) / ((2l+1)(2l+3)))     (internal.py:80-84)
that also drive the 2x2 laser rotations.
"""
import numpy as np


def cos_coupling(l, m):
    l = np.asarray(l, dtype=np.float64)
    return np.sqrt(np.clip(((l + 1.0) ** 2 - m * m) / ((2.0 * l + 1.0) * (2.0 * l + 3.0)), 0.0, None))


def gauss_legendre(n):
    return np.polynomial.legendre.leggauss(int(n))


def theta_lm(lmax, m, x):
    """[len(x), lmax+1] table of Theta_lm(x); columns l < m are zero."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((x.size, lmax + 1))
    if m > lmax:
        return out
    sin_t = np.sqrt(np.clip(1.0 - x * x, 0.0, None))
    t = np.full(x.shape, np.sqrt(0.5))
    for k in range(1, m + 1):
        t = t * sin_t * np.sqrt((2.0 * k + 1.0) / (2.0 * k))
    out[:, m] = t
    prev = np.zeros_like(x)
    cur = t
    for l in range(m, lmax):
        nxt = (x * cur - (cos_coupling(l - 1, m) if l > m else 0.0) * prev) / cos_coupling(l, m)
        out[:, l + 1] = nxt
        prev, cur = cur, nxt
    return out


def legendre_p(lmax, x):
    """[len(x), lmax+1] table of P_L(x)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty((x.size, lmax + 1))
    out[:, 0] = 1.0
    if lmax >= 1:
        out[:, 1] = x
    for L in range(1, lmax):
        out[:, L + 1] = ((2.0 * L + 1.0) * x * out[:, L] - L * out[:, L - 1]) / (L + 1.0)
    return out


class AngularTables:
    def __init__(self, lmax, n_theta, l_h, m_values):
        self.lmax, self.n_theta, self.l_h = int(lmax), int(n_theta), int(l_h)
        self.x, self.omega = gauss_legendre(self.n_theta)
        self.n_m = int(max(m_values)) + 1
        self.B = np.stack([np.sqrt(self.omega)[:, None] * theta_lm(self.lmax, m, self.x)
                           for m in range(self.n_m)])          # [n_m, n_theta, n_l]
        self.PL = legendre_p(self.l_h, self.x)                 # [n_theta, n_har]
