"""Gauss-Lobatto DVR radial grid (precompute, host FP64).

Counterpart of ``generateGaussLobattoGrid`` (src/scripts/phys/internal.py:524-556):
the N interior nodes of the (N+2)-point Lobatto rule (roots of P'_{N+1}, which
the reference reads from workingData/mesh/legendreRoots/rootsNNNNN.txt) are
mapped by r = R zeta/2 (1+x)/(1-x+zeta); the kinetic matrix is
d_ij / (r'_i r'_j) with d_ii = -Ng(Ng+1)/(3(1-x_i^2)), d_ij = -2/(x_i-x_j)^2,
Ng = N+1.  Orbitals are stored in u-units u_i = sqrt(w_i r'_i) r R(r_i), with
the Lobatto weight w_i = 2/(Ng(Ng+1) P_Ng(x_i)^2).
"""
import numpy as np


def _legendre_pair(n, x):
    """(P_n(x), P_{n-1}(x))."""
    p0 = np.ones_like(x)
    p1 = x.copy()
    for k in range(1, n):
        p0, p1 = p1, ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0)
    return p1, p0


def lobatto_nodes(N):
    """Interior Lobatto nodes, ascending: eigenvalues of the Jacobi matrix of
    the weight (1 - x^2) (Gegenbauer 3/2), refined by Newton steps on P'_{N+1}."""
    k = np.arange(1.0, N)
    beta = np.sqrt(k * (k + 2.0) / ((2.0 * k + 1.0) * (2.0 * k + 3.0)))
    x = np.linalg.eigvalsh(np.diag(beta, -1))
    n = N + 1
    for _ in range(3):
        pn, pm = _legendre_pair(n, x)
        d1 = n * (pm - x * pn) / (1.0 - x * x)
        d2 = (2.0 * x * d1 - n * (n + 1.0) * pn) / (1.0 - x * x)
        x = x - d1 / d2
    return np.sort(x)


class RadialGrid:
    def __init__(self, N, r_max, zeta=1.0):
        self.N, self.r_max, self.zeta = int(N), float(r_max), float(zeta)
        x = lobatto_nodes(self.N)
        ng = self.N + 1
        a = 0.5 * self.r_max * self.zeta
        self.x = x
        self.r = a * (1.0 + x) / (1.0 - x + self.zeta)
        self.drdx = a * (2.0 + self.zeta) / (1.0 - x + self.zeta) ** 2
        self.P = _legendre_pair(ng, x)[0]
        self.w = 2.0 / (ng * (ng + 1.0) * self.P ** 2)
        self.wr = self.w * self.drdx
        dx = x[:, None] - x[None, :]
        np.fill_diagonal(dx, 1.0)
        d = -2.0 / dx ** 2
        np.fill_diagonal(d, -ng * (ng + 1.0) / (3.0 * (1.0 - x * x)))
        self.T = d / np.outer(self.drdx, self.drdx)

    def hamiltonian(self, l, Z, v=None):
        """-T/2 + l(l+1)/2r^2 - Z/r (+ v): real symmetric radial Hamiltonian."""
        dg = 0.5 * l * (l + 1.0) / self.r ** 2 - Z / self.r
        if v is not None:
            dg = dg + v
        H = -0.5 * self.T
        H[np.diag_indices(self.N)] += dg
        return H
