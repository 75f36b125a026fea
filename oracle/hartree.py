"""Multipole Hartree potential (oracle; test infrastructure only).

Multipole moments of the density in u-units (electrons per radial node):
    q_L(r_b) = sum_j P_L(chi_j) rho_u(r_b, chi_j)
with rho_u(r_b, chi_j) = sum_k occ_k |Psi_k(r_b, j)|^2 (Psi = B u carries
sqrt(omega_j) and sqrt(w_b r'_b)); then
    V_L(r_a) = sum_b q_L(r_b) r_<^L / r_>^{L+1}
    v_H(r_a, chi_j) = sum_L P_L(chi_j) V_L(r_a).

``direct`` restates ``h_calculateHartreePotential.cuh:22-51`` (the reference's
O(N^2) sum with the b = a term included once), generalised from L <= 3 to any
L.  ``scan`` is the O(N) prefix/suffix form the kernels use, written with ratio
factors so r^L never overflows:
    A_a = A_{a-1} (r_{a-1}/r_a)^{L+1} + q_a/r_a            (b <= a)
    B_a = (r_a/r_{a+1})^L (B_{a+1} + q_{a+1}/r_{a+1})       (b > a)
    V_L(r_a) = A_a + B_a.

Pinned by: hydrogen 1s (q_0 = 4 r^2 e^{-2r} w r') gives v_H = 1/r - (1+1/r)e^{-2r};
a P_2 shell rho = r^2 e^{-r} P_2(cos theta) against its closed-form V_2; and
direct == scan to round-off.
"""
import numpy as np


def direct(q, r):
    """q: [L_H+1, N] -> V: [L_H+1, N]."""
    q = np.asarray(q, dtype=np.float64)
    nL, N = q.shape
    V = np.zeros_like(q)
    rmin = np.minimum.outer(r, r)
    rmax = np.maximum.outer(r, r)
    ratio = rmin / rmax
    for L in range(nL):
        K = ratio ** L / rmax          # K[a, b]
        V[L] = K @ q[L]
    return V


def scan(q, r):
    """Sequential in r, vectorised over L."""
    q = np.asarray(q, dtype=np.float64)
    nL, N = q.shape
    Ls = np.arange(nL, dtype=np.float64)
    V = np.zeros_like(q)
    acc = np.zeros(nL)
    for a in range(N):
        if a > 0:
            acc = acc * (r[a - 1] / r[a]) ** (Ls + 1.0)
        acc = acc + q[:, a] / r[a]
        V[:, a] = acc
    acc = np.zeros(nL)
    for a in range(N - 2, -1, -1):
        acc = (r[a] / r[a + 1]) ** Ls * (acc + q[:, a + 1] / r[a + 1])
        V[:, a] += acc
    return V
