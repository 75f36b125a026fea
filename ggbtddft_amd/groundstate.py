"""Imaginary-time FP64 ground state with Gram-Schmidt (precompute, host).

Replaces the reference's diagonalisation SCF (src/scripts/phys/groundState.py:182-457)
with an imaginary-time iteration.  One radial function per (n, l) shell, seeded
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
import ctypes
import time

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


class DeviceGroundState:
    """The same ground state computed on the GPU (ggb_ground_state, FP64).

    The device reaches the fixed point of the iteration above by
    preconditioned Rayleigh-Ritz sweeps instead of one dense eigendecomposition
    per l and iteration: per SCF iteration the damped potential update, then
    ``sweeps`` passes over the shells of each l (lower Ritz vector of
    span(u, w), w the residual H u - eps u preconditioned by the bare-nucleus
    spectral form Q0 diag(1/(|lam0 - eps| + gamma)) Q0^T).  The host computes
    only the bare-nucleus eigenpairs of each l present (the preconditioner and
    the start, independent of v).  Same attributes as GroundState; u is real.
    """

    def __init__(self, grid, atom, interacting=None, dtau=None, mix=0.3, tol=1e-11,
                 max_iter=3000, sweeps=8, gamma=2.0, device=0):
        from . import _lib
        self.grid, self.atom = grid, atom
        self.interacting = (atom.symbol != "H") if interacting is None else bool(interacting)
        N = grid.N
        shells = atom.shells
        shell_l = np.array([l for (_, l, _) in shells], dtype=np.int32)
        occ = np.array([o for (_, _, o) in shells], dtype=np.float64)
        ls = sorted(set(int(l) for l in shell_l))
        t0 = time.perf_counter()
        Q0 = np.empty((len(ls), N, N))
        lam0 = np.empty((len(ls), N))
        for g, l in enumerate(ls):
            lam0[g], Q0[g] = np.linalg.eigh(grid.hamiltonian(l, atom.Z))
        p = _lib.GsProblem(n_r=N, n_shell=len(shells), sweeps=int(sweeps), max_iter=int(max_iter),
                           interacting=int(self.interacting), device=int(device), Z=float(atom.Z),
                           mix=float(mix), gamma=float(gamma), tol=float(tol))
        u = np.zeros((len(shells), N))
        eps = np.zeros(len(shells))
        v_gs = np.zeros(N)
        rho_u = np.zeros(N)
        it = np.zeros(1, dtype=np.int32)
        conv = np.zeros(1, dtype=np.int32)
        lib = _lib.load()
        t1 = time.perf_counter()
        rc = lib.ggb_ground_state(ctypes.byref(p), _lib.iptr(shell_l), _lib.dptr(occ),
                                  _lib.dptr(np.ascontiguousarray(grid.r)),
                                  _lib.dptr(np.ascontiguousarray(grid.wr)),
                                  _lib.dptr(np.ascontiguousarray(grid.T)), _lib.dptr(Q0),
                                  _lib.dptr(lam0), _lib.dptr(u), _lib.dptr(eps), _lib.dptr(v_gs),
                                  _lib.dptr(rho_u), _lib.iptr(it), _lib.iptr(conv))
        _lib.check(rc, None, "ggb_ground_state")
        t2 = time.perf_counter()
        self.seconds = {"host_bare_eigenpairs": t1 - t0, "device_iteration": t2 - t1}
        self.u = u.astype(np.complex128)
        self.eps = eps
        self.v_gs = v_gs
        self.rho_u = rho_u
        self.iterations = int(it[0])
        self.converged = bool(conv[0])
