"""Imaginary-time FP64 ground state (oracle; test infrastructure only).

Imaginary-time iteration (the reference diagonalises instead, ``groundState.py:182-457``):
one radial function u_s per closed (n, l) shell, seeded complex128 normal start
(numpy default_rng(seed), real then imaginary part per shell, shells in atom
order), then repeated

    v_out = V_0[rho_u] + v_xc[rho]         (spherical Hartree + LDA)
    v    <- v + mix (v_out - v)            (damped potential update; the first
                                            iteration takes v = v_out)
    u_s <- Q exp(-max(lam - lam_k, 0) dtau) Q^T u_s
                                           (exact exponential of the frozen-
                                            potential radial Hamiltonian
                                            H_l = Q diag(lam) Q^T, shifted to
                                            the shell's own level lam_k, k =
                                            its rank within l, and clamped so
                                            lower levels are not amplified)
    Gram-Schmidt within each l, shells in atom order

until the orbital energies move by less than ``tol`` and v_out = v to 1e3 tol.
The damping keeps the explicit potential update stable for heavy atoms (an
undamped update at dtau = 0.2 oscillates for Kr).  The clamp matters because a
single exponential amplifies the 1s component of a 3s trial function by
e^{(lam_3s - lam_1s) dtau} (e^{100} for Ar at dtau = 1), after which Gram-Schmidt
cancels catastrophically; with the clamp Gram-Schmidt only removes O(1)
components and the fixed point is unchanged.  With the exact
exponential the fixed point is the self-consistent LDA ground state of the
discretised problem, independent of dtau.

Pinned by: hydrogen and He+ exact levels (non-interacting), orthonormality,
electron count, the virial-free check that the converged orbitals are
eigenvectors of H[v_gs] (residual), and the He LDA literature value of the
1s eigenvalue (-0.5704 Hartree for Slater+PW92 in the complete-basis limit,
checked loosely).
"""
import numpy as np

from . import hartree, xc
from .mesh import radial_hamiltonian


def spherical_potential(mesh, rho_u, Z_interacting=True):
    """v(r) = V_0 + v_xc for a spherical u-unit density rho_u(r)."""
    r = mesh["r"]
    if not Z_interacting:
        return np.zeros_like(r)
    V0 = hartree.scan(rho_u[None, :], r)[0]
    rho = rho_u / (4.0 * np.pi * mesh["wr"] * r * r)
    return V0 + xc.v_xc(rho)


def gram_schmidt(us):
    out = []
    for u in us:
        v = u.copy()
        for w in out:
            v = v - np.vdot(w, v) * w
        v = v / np.sqrt(np.vdot(v, v).real)
        out.append(v)
    return out


def ground_state(mesh, Z, shells, interacting=True, seed=0, dtau=1.0, mix=0.3,
                 tol=1e-11, max_iter=3000):
    """shells: [(n, l, occ)].  Returns dict(u=[N_shell, N] complex,
    eps=[N_shell], v_gs=[N], iterations)."""
    N = mesh["N"]
    rng = np.random.default_rng(seed)
    u = []
    for _ in shells:
        re = rng.normal(size=N)
        im = rng.normal(size=N)
        u.append(re + 1j * im)
    u = np.array(u)
    ls = sorted(set(l for (_, l, _) in shells))
    groups = {l: [s for s, (_, ll, _) in enumerate(shells) if ll == l] for l in ls}
    occ = np.array([o for (_, _, o) in shells])
    for l, idx in groups.items():
        u[idx] = gram_schmidt(list(u[idx]))
    eps_old = np.full(len(shells), np.inf)
    v = None
    it = 0
    for it in range(1, max_iter + 1):
        rho_u = np.einsum("s,si->i", occ, np.abs(u) ** 2)
        v_out = spherical_potential(mesh, rho_u, interacting)
        v = v_out if v is None else v + mix * (v_out - v)
        eps = np.zeros(len(shells))
        for l, idx in groups.items():
            H = radial_hamiltonian(mesh, l, Z, v)
            lam, Q = np.linalg.eigh(H)
            newu = []
            for rank, s in enumerate(idx):
                decay = np.exp(-np.maximum(lam - lam[rank], 0.0) * dtau)
                newu.append(Q @ (decay * (Q.T @ u[s])))
            newu = gram_schmidt(newu)
            for s, w in zip(idx, newu):
                u[s] = w
                eps[s] = np.vdot(w, H @ w).real
        if np.max(np.abs(eps - eps_old)) < tol and np.max(np.abs(v_out - v)) < 1e3 * tol:
            break
        eps_old = eps
    rho_u = np.einsum("s,si->i", occ, np.abs(u) ** 2)
    v_gs = spherical_potential(mesh, rho_u, interacting)
    return dict(u=u, eps=eps, v_gs=v_gs, iterations=it, rho_u=rho_u)
