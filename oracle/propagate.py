"""Real-time split-operator propagation (oracle; test infrastructure only).

State: psi[l, k, i] complex128 (l = 0..Lmax, orbital k, radial node i) in the
reference's u-units (u = sqrt(w r') * r R(r)), zero for l < m_k; the layout of
``ksoSetTD.d_x[l][i + k N]`` in ``main.cu:363-391``.

One step n (t_n -> t_n + dt) is the palindrome
    S = Vh . Le . Lo . K . Lo . Le . Vh          (right-most acts first)
* Vh: theta space, Psi = B u, Psi *= exp(-i dV dt/2) * mask, u = B^T Psi.
  dV(r, theta) = v_H + v_xc - v_gs is built from psi_n (explicit update, as the
  reference recomputes potentials after each full step, h_timePropagation.cuh:97-101);
  mask = exp(-W_cap dt/2).
* Le, Lo: length-gauge laser r E(t_n + dt/2) cos(theta) for dt/2, as exact 2x2
  rotations on the (l, l+1) blocks of the tridiagonal cos(theta) matrix,
  even blocks (l - m even) and odd blocks; phase sign e^{-i eta} as
  ``couplingP.cuh:99-104``.
* K: exp(-i H_l dt) with H_l = -T/2 + l(l+1)/2r^2 - Z/r + v_gs = Q diag(lam) Q^T,
  applied as Q (e^{-i lam dt} (Q^T u)) (the reference's per-l
  ``cublasZgemmBatched``, h_timePropagation.cuh:17-26, with a CN W_l),
  restricted to the eigenpairs with lam <= E_cut = ecut_factor * pi / dt.
  Eigencomponents above E_cut are dropped: their phase turns by more than pi
  per step, so the splitting cannot resolve them, and with the theta-dependent
  LDA they resonate with the pointwise potential at the innermost radial
  nodes (high l, lam ~ l(l+1)/2r^2 ~ 1e6) and grow without bound (He, 1e14
  W/cm^2: dV reaches 10 Hartree at t ~ 140 for dt = 0.05 and for dt = 0.01
  alike).  With the cut the run is stable and the norm moves by 2e-10.
  The retained columns Q[:, :nkeep_l] are what the kernels stream.
After the step: Psi = B psi_{n+1}; potentials for step n+1 and the observables
    dip_k  = sum r_i chi_j |Psi_k|^2      (<r cos theta>, dipoleMoment.cuh)
    acc_k  = sum chi_j |Psi_k|^2 / r_i^2  (<cos theta / r^2>, dipoleAcceleration.cuh)
    norm_k = sum |Psi_k|^2
all from that same Psi.

CAP: Manolopoulos form, restating ``internal.py:460-475`` (getABCR, ABC_TYPE 1)
with R_MASK = r_max / 1.5 (init.py:284), applied once (the reference adds it
to both v0 and the propagator, an effective -2iW).
"""
import numpy as np

from . import angular, atoms, groundstate, hartree, mesh as meshmod, xc


def manolopoulos_cap(r, r_mask, r_max):
    delta = 0.20
    c = 2.62206
    a = 1.0 - 16.0 / c ** 3
    b = (1.0 - 17.0 / c ** 3) / c ** 2
    kmin = c / (delta * 2.0 * (r_max - r_mask))
    emin = 0.5 * kmin ** 2
    W = np.zeros_like(r)
    i0 = int(np.argmin(np.abs(r - r_mask)))
    rabc = r[i0:] - r_mask
    x = 2.0 * delta * kmin * rabc
    y = a * x - b * x ** 3 + 4.0 / (c - x) ** 2 - 4.0 / (c + x) ** 2
    W[i0:] = emin * y * rabc
    return W


PROPAGATORS = ("exp", "cn", "cn2", "cn3")


def kinetic_phase(propagator, lam, dt):
    """Factor of the kinetic propagator on an eigenpair of H_l (see
    System.kinetic_phase)."""
    lam = np.asarray(lam, dtype=np.float64)
    if propagator == "exp":
        return np.exp(-1j * dt * lam)
    n = {"cn": 1, "cn2": 2, "cn3": 3}[propagator]
    x = 0.5 * dt * lam
    p = np.zeros_like(x, dtype=np.complex128)
    term = np.ones_like(x, dtype=np.complex128)
    for k in range(n + 1):
        p = p + term
        term = term * (1j * x) / (k + 1)
    return np.conj(p) / p


class System:
    """Static data of one atom on one grid (everything a run needs but the field)."""

    def __init__(self, atom, N, r_max, lmax, dt, zeta=1.0, n_theta=None, l_h=None,
                 seed=0, cap=True, interacting=None, gs=None, ecut_factor=1.0, propagator="exp"):
        self.atom = atoms.canonical(atom)
        self.Z = atoms.nuclear_charge(self.atom)
        self.N, self.r_max, self.lmax, self.dt, self.zeta = N, r_max, lmax, dt, zeta
        self.nl = lmax + 1
        self.l_h = 2 * lmax if l_h is None else l_h
        self.n_theta = (2 * lmax + 1) if n_theta is None else n_theta
        self.interacting = (self.atom != "H") if interacting is None else interacting
        self.mesh = meshmod.lobatto_mesh(N, r_max, zeta)
        m = self.mesh
        self.r, self.wr = m["r"], m["wr"]
        self.shells = atoms.shells(self.atom)
        self.orbs = atoms.td_orbitals(self.atom)
        self.m = np.array([o[2] for o in self.orbs])
        self.occ = np.array([o[3] for o in self.orbs])
        self.norb = len(self.orbs)
        chi_x, omega_x = angular.theta_grid(self.n_theta, extended=True)
        self.chi, self.omega = chi_x.astype(np.float64), omega_x.astype(np.float64)
        self.B = {mm: angular.transform_matrix(lmax, mm, chi_x, omega_x)
                  for mm in sorted(set(self.m))}
        self.PL = angular.legendre_table(self.l_h, chi_x)
        if gs is None:
            gs = groundstate.ground_state(m, self.Z, self.shells, self.interacting, seed=seed)
        self.gs = gs
        self.v_gs = gs["v_gs"]
        self.ecut = ecut_factor * np.pi / dt
        if propagator not in PROPAGATORS:
            raise ValueError("propagator: one of %s" % (PROPAGATORS,))
        self.propagator = propagator
        self.Q, self.lam, self.nkeep = [], [], []
        for l in range(self.nl):
            H = meshmod.radial_hamiltonian(m, l, self.Z, self.v_gs)
            lam, Q = np.linalg.eigh(H)
            nk = int(np.searchsorted(lam, self.ecut, side="right"))
            self.Q.append(Q[:, :nk])
            self.lam.append(lam[:nk])
            self.nkeep.append(nk)
        W = manolopoulos_cap(self.r, r_max / 1.5, r_max) if cap else np.zeros(N)
        self.W = W
        self.mask = np.exp(-0.5 * dt * W)
        self.psi0 = np.zeros((self.nl, self.norb, N), dtype=np.complex128)
        for k, (n, l, mm, occ, s) in enumerate(self.orbs):
            self.psi0[l, k] = gs["u"][s]

    # ---- potentials and observables -------------------------------------
    def theta_state(self, psi):
        """Psi[k, i, j] = sum_l B^{m_k}[j, l] psi[l, k, i]."""
        out = np.empty((self.norb, self.N, self.n_theta), dtype=np.complex128)
        for k in range(self.norb):
            out[k] = psi[:, k, :].T @ self.B[self.m[k]].T
        return out

    def potential(self, Psi, rho_frozen_u=None):
        """dV[i, j] from the theta-space state."""
        if not self.interacting:
            return np.zeros((self.N, self.n_theta))
        rho_u = np.einsum("k,kij->ij", self.occ, np.abs(Psi) ** 2)
        if rho_frozen_u is not None:
            rho_u = rho_u + 0.5 * rho_frozen_u[:, None] * self.omega[None, :]
        rho = rho_u / (2.0 * np.pi * self.omega[None, :] * (self.wr * self.r * self.r)[:, None])
        vxc = xc.v_xc(rho)
        q = (rho_u @ self.PL).T                      # [L, i]
        V = hartree.scan(q, self.r)                  # [L, i]
        vH = (self.PL @ V).T                         # [i, j]
        return vH + vxc - self.v_gs[:, None]

    def observables(self, Psi):
        P2 = np.abs(Psi) ** 2
        dip = np.einsum("kij,i,j->k", P2, self.r, self.chi)
        acc = np.einsum("kij,i,j->k", P2, 1.0 / self.r ** 2, self.chi)
        nrm = P2.sum(axis=(1, 2))
        return dip, acc, nrm

    def channel_dipoles(self, psi):
        """chan[k, l] = <x_l | r (c_u(l) x_{l+1} + c_d(l) x_{l-1})> of orbital k
        (complex, conjugate on the left), the per-channel dipole of
        ``dipoleMoment.cuh:1-47`` reduced by ``cublasZdotc`` (:71);
        c_u(l) = c(l, m), c_d(l) = c(l - 1, m) (angularCoefficients.cuh d_cu,
        d_cd), zero for l < m.  sum_l Re chan[k, l] = dip_k."""
        out = np.zeros((self.norb, self.nl), dtype=np.complex128)
        for k in range(self.norb):
            mm = self.m[k]
            for l in range(mm, self.nl):
                z = np.zeros(self.N, dtype=np.complex128)
                if l + 1 < self.nl:
                    z += angular.c_lm(l, mm) * psi[l + 1, k]
                if l - 1 >= mm:
                    z += angular.c_lm(l - 1, mm) * psi[l - 1, k]
                out[k, l] = np.vdot(psi[l, k], self.r * z)
        return out

    # ---- sub-steps ---------------------------------------------------------
    def v_half(self, psi, dV):
        ph = np.exp(-0.5j * self.dt * dV) * self.mask[:, None]
        for k in range(self.norb):
            B = self.B[self.m[k]]
            Psi = psi[:, k, :].T @ B.T
            psi[:, k, :] = ((Psi * ph) @ B).T

    def laser(self, psi, E, parity):
        a = 0.5 * self.dt * E * self.r
        for k in range(self.norb):
            mm = self.m[k]
            even, odd = angular.pair_lists(self.lmax, mm)
            for l in (even if parity == 0 else odd):
                th = a * angular.c_lm(l, mm)
                cs, sn = np.cos(th), np.sin(th)
                x0 = psi[l, k].copy()
                x1 = psi[l + 1, k].copy()
                psi[l, k] = cs * x0 - 1j * sn * x1
                psi[l + 1, k] = -1j * sn * x0 + cs * x1

    def kinetic_phase(self, lam):
        """Eigenvalue factor of the kinetic propagator: e^{-i lam dt} ("exp",
        the reference's EIG/expm), or for its Crank-Nicolson forms CN, CN2,
        CN3 (propagatorsR, internal.py:252-318: W = Wp^{-1} Wm with Wp, Wm the
        truncated series of e^{+-i dt H / 2}, the same eigenvectors)
        conj(p(x)) / p(x), p(x) = sum_{k<=n} (i x)^k / k!, x = lam dt / 2."""
        return kinetic_phase(self.propagator, lam, self.dt)

    def kinetic(self, psi):
        for l in range(self.nl):
            ph = self.kinetic_phase(self.lam[l])
            Y = self.Q[l].T @ psi[l].T               # [nkeep, k]
            psi[l] = (self.Q[l] @ (ph[:, None] * Y)).T

    def step(self, psi, dV, E):
        self.v_half(psi, dV)
        self.laser(psi, E, 0)
        self.laser(psi, E, 1)
        self.kinetic(psi)
        self.laser(psi, E, 1)
        self.laser(psi, E, 0)
        self.v_half(psi, dV)

    def run(self, E_mid, psi=None, nsteps=None, channels=False):
        """Propagate; returns dict(dip, acc, norm [nsteps, norb], psi) and,
        with channels, chan [nsteps, norb, nl] (channel_dipoles after each step)."""
        psi = self.psi0.copy() if psi is None else psi.copy()
        nsteps = len(E_mid) if nsteps is None else nsteps
        Psi = self.theta_state(psi)
        dV = self.potential(Psi)
        dip = np.empty((nsteps, self.norb))
        acc = np.empty((nsteps, self.norb))
        nrm = np.empty((nsteps, self.norb))
        chan = np.empty((nsteps, self.norb, self.nl), dtype=np.complex128) if channels else None
        for n in range(nsteps):
            self.step(psi, dV, E_mid[n])
            Psi = self.theta_state(psi)
            dip[n], acc[n], nrm[n] = self.observables(Psi)
            if channels:
                chan[n] = self.channel_dipoles(psi)
            dV = self.potential(Psi)
        out = dict(dip=dip, acc=acc, norm=nrm, psi=psi)
        if channels:
            out["chan"] = chan
        return out

    def total_acceleration(self, acc, norm, E_end):
        """a(t) = -Z sum_k occ_k <cos/r^2>_k - E(t) sum_k occ_k |psi_k|^2
        (Ehrenfest; Hartree and LDA forces sum to zero over all electrons)."""
        return -self.Z * acc @ self.occ - E_end * (norm @ self.occ)
