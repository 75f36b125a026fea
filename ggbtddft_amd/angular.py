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


_LD = np.longdouble


def cos_coupling(l, m):
    l = np.asarray(l, dtype=np.float64)
    return np.sqrt(np.clip(((l + 1.0) ** 2 - m * m) / ((2.0 * l + 1.0) * (2.0 * l + 3.0)), 0.0, None))


def _legendre_ld(n, x):
    """(P_n, P'_n) in extended precision."""
    p0 = np.ones_like(x)
    p1 = x.copy()
    for k in range(1, n):
        p0, p1 = p1, ((2 * k + 1) * x * p1 - k * p0) / (k + 1)
    return p1, n * (x * p1 - p0) / (x * x - 1)


def gauss_legendre(n):
    """Nodes and weights built in extended precision (Newton-polished nodes,
    w = 2/((1-x^2) P_n'(x)^2)) and rounded once, so the transforms below are
    orthonormal to ~1e-16 without a systematic bias (a double-precision
    construction leaves B^T B - I ~ 3e-13 on the diagonal at Lmax = 64, which
    shows up as a steady norm drift)."""
    n = int(n)
    x = np.polynomial.legendre.leggauss(n)[0].astype(_LD)
    for _ in range(3):
        p, dp = _legendre_ld(n, x)
        x = x - p / dp
    p, dp = _legendre_ld(n, x)
    w = 2 / ((1 - x * x) * dp * dp)
    return x.astype(np.float64), w.astype(np.float64), x, w


def _theta_lm_ld(lmax, m, x):
    out = np.zeros((x.size, lmax + 1), dtype=_LD)
    if m > lmax:
        return out
    sin_t = np.sqrt(np.clip(1 - x * x, 0, None))
    t = np.full(x.shape, np.sqrt(_LD(0.5)), dtype=_LD)
    for k in range(1, m + 1):
        t = t * sin_t * np.sqrt(_LD(2 * k + 1) / _LD(2 * k))
    out[:, m] = t
    cc = lambda l: np.sqrt(_LD((l + 1) ** 2 - m * m) / _LD((2 * l + 1) * (2 * l + 3)))
    prev = np.zeros_like(x)
    cur = t
    for l in range(m, lmax):
        nxt = (x * cur - (cc(l - 1) if l > m else 0) * prev) / cc(l)
        out[:, l + 1] = nxt
        prev, cur = cur, nxt
    return out


def theta_lm(lmax, m, x):
    """[len(x), lmax+1] table of Theta_lm(x); columns l < m are zero."""
    return _theta_lm_ld(lmax, m, np.asarray(x, dtype=_LD)).astype(np.float64)


def legendre_p(lmax, x):
    """[len(x), lmax+1] table of P_L(x)."""
    x = np.asarray(x, dtype=_LD)
    out = np.empty((x.size, lmax + 1), dtype=_LD)
    out[:, 0] = 1
    if lmax >= 1:
        out[:, 1] = x
    for L in range(1, lmax):
        out[:, L + 1] = ((2 * L + 1) * x * out[:, L] - L * out[:, L - 1]) / (L + 1)
    return out.astype(np.float64)


class AngularTables:
    def __init__(self, lmax, n_theta, l_h, m_values):
        self.lmax, self.n_theta, self.l_h = int(lmax), int(n_theta), int(l_h)
        self.x, self.omega, xl, wl = gauss_legendre(self.n_theta)
        self.n_m = int(max(m_values)) + 1
        self.B = np.stack([(np.sqrt(wl)[:, None] * _theta_lm_ld(self.lmax, m, xl)).astype(np.float64)
                           for m in range(self.n_m)])          # [n_m, n_theta, n_l]
        self.PL = legendre_p(self.l_h, xl)                     # [n_theta, n_har]
