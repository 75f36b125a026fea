"""Imaginary-time FP64 ground state with Gram-Schmidt (precompute, host).

Replaces the reference's diagonalisation SCF (src/scripts/phys/groundState.py:182-457)
with the north-star method.  One radial function per (n, l) shell, seeded
complex128 normal start (default_rng(seed): real then imaginary N-vector per
shell, shells in atom order).  Each iteration:
    v_out = V_0[rho_u] + v_xc[rho]          spherical Hartree + LDA
    v    <- v + mix (v_out - v)            damped explicit update
    u_s  <- Q exp(-max(lam - lam_k, 0) dtau) Q^T u_s
            (H_l[v] = Q diag(lam) Q^T; k = rank of shell s within its l; the
             clamp stops deeper levels from being amplified, which would make
             the subsequent Gram-Schmidt cancel catastrophically)
    Gram-Schmidt within each l in shell order
until the shell energies change by less than tol and v_out = v to 1e3 tol.
"""
import numpy as np

from . import lda


def hartree_l0(q, r):
    """V_0(r_a) = sum_b q_b / max(r_a, r_b) for a u-unit density q."""
    inner = np.cumsum(q) / r
    outer = np.concatenate([np.cumsum((q / r)[::-1])[::-1][1:], [0.0]])
    return inner + outer


def spherical_potential(grid, rho_u, interacting=True):
    if not interacting:
        return np.zeros(grid.N)
    rho = rho_u / (4.0 * np.pi * grid.wr * grid.r ** 2)
    return hartree_l0(rho_u, grid.r) + lda.vxc(rho)


def _orthonormalise(vectors):
    basis = []
    for v in vectors:
        w = v.copy()
        for b in basis:
            w -= np.vdot(b, w) * b
        basis.append(w / np.sqrt(np.vdot(w, w).real))
    return basis


class GroundState:
    def __init__(self, grid, atom, interacting=None, seed=0, dtau=1.0, mix=0.3, tol=1e-11,
                 max_iter=3000):
        self.grid, self.atom = grid, atom
        self.interacting = (atom.symbol != "H") if interacting is None else bool(interacting)
        N = grid.N
        rng = np.random.default_rng(seed)
        shells = atom.shells
        u = np.empty((len(shells), N), dtype=np.complex128)
        for s in range(len(shells)):
            re = rng.normal(size=N)
            im = rng.normal(size=N)
            u[s] = re + 1j * im
        occ = np.array([o for (_, _, o) in shells])
        by_l = {}
        for s, (_, l, _) in enumerate(shells):
            by_l.setdefault(l, []).append(s)
        for l, idx in by_l.items():
            u[idx] = _orthonormalise(list(u[idx]))
        eps = np.zeros(len(shells))
        eps_prev = np.full(len(shells), np.inf)
        v = None
        it = 0
        for it in range(1, max_iter + 1):
            rho_u = occ @ (np.abs(u) ** 2)
            v_out = spherical_potential(grid, rho_u, self.interacting)
            v = v_out if v is None else v + mix * (v_out - v)
            for l, idx in by_l.items():
                H = grid.hamiltonian(l, atom.Z, v)
                lam, Q = np.linalg.eigh(H)
                prop = []
                for rank, s in enumerate(idx):
                    g = np.exp(-np.maximum(lam - lam[rank], 0.0) * dtau)
                    prop.append(Q @ (g * (Q.T @ u[s])))
                for s, w in zip(idx, _orthonormalise(prop)):
                    u[s] = w
                    eps[s] = np.vdot(w, H @ w).real
            if np.max(np.abs(eps - eps_prev)) < tol and np.max(np.abs(v_out - v)) < 1e3 * tol:
                break
            eps_prev = eps.copy()
        self.iterations = it
        self.converged = it < max_iter
        self.u = u
        self.eps = eps
        self.rho_u = occ @ (np.abs(u) ** 2)
        self.v_gs = spherical_potential(grid, self.rho_u, self.interacting)
