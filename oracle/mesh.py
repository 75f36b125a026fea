"""Gauss-Lobatto DVR radial mesh (oracle; test infrastructure only).

Restates ``src/scripts/phys/internal.py:524-556`` (generateGaussLobattoGrid):

* interior Lobatto nodes x_i, i = 1..N: the roots of P'_{N+1}(x)
  (the reference reads them from ``workingData/mesh/legendreRoots``);
* mapping r = R*zeta/2 * (1+x)/(1-x+zeta)          (internal.py:530)
* dr/dx  = R*zeta/2 * (2+zeta)/(1-x+zeta)^2        (internal.py:532)
* d_ii = -Ng(Ng+1)/(3(1-x_i^2)), d_ij = -2/(x_i-x_j)^2, Ng = N+1
  and T_ij = d_ij/(r'_i r'_j)                       (internal.py:540-548)
* P = P_Ng(x)                                       (internal.py:554)

The Lobatto weights w_i = 2/(Ng(Ng+1) P_Ng(x_i)^2) are implied by the
reference's "u-unit" density scaling Ng(Ng+1)/2 * P^2 / r'
(``h_calculateExchangeCorrelationPotential.cuh:30``, ``groundState.py:306``),
which is 1/(w_i r'_i).

Pinned by: the reference's root files (committed under tests/golden/roots),
hydrogen's -1/(2n^2) levels per l (closed form), and the quadrature identity
sum_i w_i r'_i f(r_i) = int_0^R f(r) dr for smooth f vanishing at the ends.
"""
import numpy as np


def legendre_p_and_dp(n, x):
    """P_n(x) and P_{n-1}(x) by the three-term recurrence (textbook)."""
    x = np.asarray(x, dtype=np.float64)
    p_prev = np.ones_like(x)
    p = x.copy()
    if n == 0:
        return p_prev, np.zeros_like(x)
    for k in range(1, n):
        p_prev, p = p, ((2 * k + 1) * x * p - k * p_prev) / (k + 1)
    return p, p_prev


def gl_interior_nodes(N):
    """Roots of P'_{N+1}: Golub-Welsch on the Gegenbauer(3/2) Jacobi matrix,
    polished by Newton on P'_{N+1} using the Legendre ODE for P''."""
    n = np.arange(1, N, dtype=np.float64)
    off = np.sqrt(n * (n + 2.0) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)))
    J = np.diag(off, 1) + np.diag(off, -1)
    x = np.sort(np.linalg.eigvalsh(J))
    m = N + 1
    for _ in range(3):
        p, pm1 = legendre_p_and_dp(m, x)
        dp = m * (x * p - pm1) / (x * x - 1.0)
        d2p = (2.0 * x * dp - m * (m + 1.0) * p) / (1.0 - x * x)
        x = x - dp / d2p
    return x


def lobatto_mesh(N, R, zeta):
    """Return dict with x, r, dr (= dr/dx), w (Lobatto weights), wr = w*dr,
    P (= P_{N+1}(x)), Ng and the kinetic matrix T (d^2/dr^2 in u-units)."""
    x = gl_interior_nodes(N)
    Ng = N + 1
    r = R * zeta / 2.0 * (x + 1.0) / (1.0 - x + zeta)
    dr = R * zeta / 2.0 * (2.0 + zeta) / (1.0 - x + zeta) ** 2
    P, _ = legendre_p_and_dp(Ng, x)
    w = 2.0 / (Ng * (Ng + 1.0) * P * P)
    d = np.empty((N, N))
    for i in range(N):
        for j in range(N):
            if i == j:
                d[i, j] = -Ng * (Ng + 1.0) / (3.0 * (1.0 - x[i] ** 2))
            else:
                d[i, j] = -2.0 / (x[i] - x[j]) ** 2
    T = d / np.outer(dr, dr)
    return dict(N=N, R=R, zeta=zeta, Ng=Ng, x=x, r=r, dr=dr, w=w, wr=w * dr,
                P=P, T=T)


def radial_hamiltonian(mesh, l, Z, v=None):
    """H_l = -T/2 + l(l+1)/(2r^2) - Z/r (+ v), real symmetric.
    Same operator as groundState.py:243-247 without the CAP term."""
    r = mesh["r"]
    diag = l * (l + 1.0) / (2.0 * r * r) - Z / r
    if v is not None:
        diag = diag + v
    return -0.5 * mesh["T"] + np.diag(diag)
