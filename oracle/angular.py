"""Angular machinery (oracle; test infrastructure only).

Basis: for an orbital of magnetic number m >= 0,
    psi(r, theta, phi) = sum_{l=m}^{Lmax} u_l(r)/(r sqrt(w r')) * Y_lm(theta, phi),
    Y_lm = Theta_lm(cos theta) e^{i m phi} / sqrt(2 pi),
    int_{-1}^{1} Theta_lm Theta_l'm dx = delta_ll'.

* c(l, m) = sqrt(((l+1)^2 - m^2) / ((2l+1)(2l+3))) = <l m|cos theta|l+1 m>;
  restates ``internal.py:80-84`` (c) and ``angularCoefficients.cuh`` (d_c).
* Theta_lm from Theta_mm and the three-term recurrence
  x Theta_lm = c(l,m) Theta_{l+1,m} + c(l-1,m) Theta_{l-1,m}.
  (Condon-Shortley phase dropped; it is constant for fixed m.)
* theta grid: Gauss-Legendre nodes chi_j = cos(theta_j) and weights omega_j.
* transform matrix per m: B[j, l] = sqrt(omega_j) Theta_lm(chi_j), zero for
  l < m.  Forward Psi = B u, back u = B^T Psi.  B^T B = I when the rule is
  exact for degree 2 Lmax, i.e. N_theta >= Lmax + 1.
* Legendre table PL[j, L] = P_L(chi_j), L = 0..L_H.

Pinned by: scipy.special.lpmv (textbook associated Legendre functions),
orthonormality, and the reference's m=0 Clebsch-Gordan tables
(``workingData/stateParameters/clebschGordan/lo00{0,2}.txt``) through
<l1 0|P_L|l2 0> = sqrt((2l1+1)(2l2+1)) <l1 0 l2 0|L 0>^2 / (2L+1).
"""
import numpy as np


def c_lm(l, m):
    """<l m|cos theta|l+1 m> (internal.py:80-84); zero outside l >= m >= 0."""
    l = np.asarray(l, dtype=np.float64)
    val = np.sqrt(np.maximum(((l + 1.0) ** 2 - m * m) / ((2 * l + 1.0) * (2 * l + 3.0)), 0.0))
    return np.where(l >= m, val, 0.0)


EXT = np.longdouble   # tables are built in extended precision, rounded once


def theta_grid(n_theta, extended=False):
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1]: numpy's
    leggauss refined by Newton steps on P_n in extended precision, weights
    2/((1-x^2) P_n'^2).  Rounded to double unless extended=True."""
    chi = np.polynomial.legendre.leggauss(n_theta)[0].astype(EXT)
    for it in range(4):
        pm, p = np.ones_like(chi), chi.copy()
        for k in range(1, n_theta):
            pm, p = p, ((2 * k + 1) * chi * p - k * pm) / (k + 1)
        dp = n_theta * (chi * p - pm) / (chi * chi - 1)
        if it < 3:
            chi = chi - p / dp
    omega = 2 / ((1 - chi * chi) * dp * dp)
    if extended:
        return chi, omega
    return chi.astype(np.float64), omega.astype(np.float64)


def assoc_legendre_normalised(lmax, m, x):
    """Theta_lm(x) for l = 0..lmax as an array [len(x), lmax+1] (zero for l<m),
    computed in the precision of x (pass extended-precision nodes to get
    tables rounded once)."""
    x = np.asarray(x)
    if x.dtype != EXT:
        x = x.astype(np.float64)
    out = np.zeros((x.size, lmax + 1), dtype=x.dtype)
    if m > lmax:
        return out
    # Theta_mm = sqrt((2m+1)/2 * (2m)!/(2^m m!)^2 ... ) (1-x^2)^{m/2}, built
    # iteratively: Theta_00 = 1/sqrt(2), Theta_kk = sqrt((2k+1)/(2k)) s Theta_{k-1,k-1}
    one = x.dtype.type(1)
    s = np.sqrt(np.maximum(one - x * x, 0))
    t = np.full_like(x, one / np.sqrt(x.dtype.type(2)))
    for k in range(1, m + 1):
        t = np.sqrt(x.dtype.type(2 * k + 1) / x.dtype.type(2 * k)) * s * t
    out[:, m] = t
    c = lambda l: np.sqrt(x.dtype.type((l + 1) ** 2 - m * m) / x.dtype.type((2 * l + 1) * (2 * l + 3)))
    if m + 1 <= lmax:
        out[:, m + 1] = x * out[:, m] / c(m)
    for l in range(m + 1, lmax):
        out[:, l + 1] = (x * out[:, l] - c(l - 1) * out[:, l - 1]) / c(l)
    return out


def transform_matrix(lmax, m, chi, omega):
    """B[j, l] = sqrt(omega_j) Theta_lm(chi_j), in double.  With extended-
    precision chi/omega (theta_grid(n, extended=True)) the table is rounded
    once: B^T B = I to ~1e-16 with no systematic bias."""
    return (np.sqrt(omega)[:, None] * assoc_legendre_normalised(lmax, m, chi)).astype(np.float64)


def legendre_table(L_H, x):
    """PL[j, L] = P_L(x_j), L = 0..L_H (in the precision of x, returned in double)."""
    x = np.asarray(x)
    if x.dtype != EXT:
        x = x.astype(np.float64)
    out = np.zeros((x.size, L_H + 1), dtype=x.dtype)
    out[:, 0] = 1
    if L_H >= 1:
        out[:, 1] = x
    for k in range(1, L_H):
        out[:, k + 1] = ((2 * k + 1) * x * out[:, k] - k * out[:, k - 1]) / (k + 1)
    return out.astype(np.float64)


def pair_lists(lmax, m):
    """(even, odd) lists of l for the 2x2 laser blocks (l, l+1), l >= m:
    even blocks start at l - m even, odd blocks at l - m odd."""
    even = [l for l in range(m, lmax) if (l - m) % 2 == 0]
    odd = [l for l in range(m, lmax) if (l - m) % 2 == 1]
    return even, odd
