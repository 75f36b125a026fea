"""Static data of one atom on one grid: the precompute stage.

Counterpart of src/scripts/init.py:304-337 (mesh, absorber, ground state,
radial propagators, angular tables) with this propagator's choices:
* kinetic propagator exp(-i H_l dt), H_l = -T/2 + l(l+1)/2r^2 - Z/r + v_gs,
  kept in spectral form Q_l diag(e^{-i lam dt}) Q_l^T on the eigenpairs with
  lam <= E_cut = ecut_factor * pi / dt (components above turn by more than pi
  per step, are not resolved by the splitting, and with the theta-dependent
  LDA they resonate at the innermost nodes; dropping them keeps the scheme
  stable and makes the streamed tables smaller);
* absorber: Manolopoulos CAP of internal.py:460-475 (getABCR, ABC_TYPE 1) on
  [r_max/1.5, r_max] (R_MAX = 1.5 R_MASK, init.py:284), applied once per step
  as the mask exp(-W dt/2) in each potential half step;
* theta grid of 2 Lmax + 1 Gauss-Legendre nodes and Hartree multipoles up to
  L_H = 2 Lmax by default (exact multipoles of a density built from l <= Lmax).
"""
import math
import os

import numpy as np

from .angular import AngularTables
from .atoms import Atom
from .grid import RadialGrid
from .groundstate import DeviceGroundState, GroundState


PROPAGATORS = ("exp", "cn", "cn2", "cn3")


def manolopoulos_cap(r, r_mask, r_max):
    delta, c = 0.20, 2.62206
    a = 1.0 - 16.0 / c ** 3
    b = (1.0 - 17.0 / c ** 3) / c ** 2
    kmin = c / (2.0 * delta * (r_max - r_mask))
    emin = 0.5 * kmin * kmin
    W = np.zeros_like(r)
    i0 = int(np.argmin(np.abs(r - r_mask)))
    d = r[i0:] - r_mask
    x = 2.0 * delta * kmin * d
    W[i0:] = emin * d * (a * x - b * x ** 3 + 4.0 / (c - x) ** 2 - 4.0 / (c + x) ** 2)
    return W


def _blas_threads(n):
    try:
        from threadpoolctl import threadpool_limits
        return threadpool_limits(limits=n)
    except Exception:  # threadpoolctl absent: leave the BLAS as it is
        import contextlib
        return contextlib.nullcontext()


def _round16(n):
    return ((int(n) + 15) // 16) * 16


class AtomSystem:
    def __init__(self, atom, N, r_max, lmax, dt, zeta=1.0, n_theta=None, l_h=None, seed=0,
                 cap=True, interacting=None, ecut_factor=1.0, ground_state=None, gs_device=0,
                 tables="host", propagator="exp"):
        self.atom = atom if isinstance(atom, Atom) else Atom(atom)
        self.N, self.r_max, self.lmax, self.dt = int(N), float(r_max), int(lmax), float(dt)
        self.n_l = self.lmax + 1
        self.l_h = 2 * self.lmax if l_h is None else int(l_h)
        self.n_har = self.l_h + 1
        self.n_theta = 2 * self.lmax + 1 if n_theta is None else int(n_theta)
        self.grid = RadialGrid(self.N, self.r_max, zeta)
        orbs = self.atom.orbitals
        self.n_orb = len(orbs)
        self.m = np.array([o[2] for o in orbs], dtype=np.int32)
        self.occ = np.array([o[3] for o in orbs], dtype=np.float64)
        self.angular = AngularTables(self.lmax, self.n_theta, self.l_h, self.m)
        # ground_state: None (host iteration), "device" (ggb_ground_state on
        # gs_device) or a computed ground state of this grid and atom
        if ground_state is None:
            self.gs = GroundState(self.grid, self.atom, interacting=interacting, seed=seed)
        elif isinstance(ground_state, str) and ground_state == "device":
            self.gs = DeviceGroundState(self.grid, self.atom, interacting=interacting,
                                        device=gs_device)
        else:
            self.gs = ground_state
        self.interacting = self.gs.interacting
        self.v_gs = self.gs.v_gs
        self.ecut = ecut_factor * math.pi / self.dt
        # kinetic propagator: "exp" = exp(-i H_l dt) (the reference's EIG / expm);
        # "cn", "cn2", "cn3" = its Crank-Nicolson forms W_l = Wp^{-1} Wm with Wp,
        # Wm the truncated series of exp(+-i dt H_l / 2) to order 1, 2, 3
        # (propagatorsR, internal.py:252-318).  All are diagonal in the
        # eigenvectors of H_l with factor conj(p(x))/p(x) = e^{-i lam' dt},
        # p(x) = sum_{k<=n} (i x)^k / k!, x = lam dt/2, lam' = 2 arg(p(x)) / dt:
        # the device applies e^{-i lam' dt}, so only the eigenvalues handed over change
        if propagator not in PROPAGATORS:
            raise ValueError("propagator: one of %s" % (PROPAGATORS,))
        self.propagator = propagator
        # kinetic propagator tables: "host" (numpy eigh per l), "device"
        # (ggb_propagator_tables: one batched eigendecomposition on GPU
        # gs_device), or None (the caller runs build_tables later)
        if tables is not None:
            self.build_tables(tables, gs_device)
        self.W = manolopoulos_cap(self.grid.r, self.r_max / 1.5, self.r_max) if cap else np.zeros(self.N)
        self.mask = np.exp(-0.5 * self.dt * self.W)
        self.psi0 = np.zeros((self.n_l, self.n_orb, self.N), dtype=np.complex128)
        for k, (n, l, m, occ, s) in enumerate(orbs):
            self.psi0[l, k] = self.gs.u[s]

    def build_tables(self, where="host", device=0):
        """Eigenpairs of every H_l with lambda <= E_cut: Q, lam, nkeep, nk_max."""
        if where == "device":
            self._tables_device(device)
        elif where == "host":
            self._tables_host()
        else:
            raise ValueError("tables: 'host' or 'device'")

    def _tables_host(self):
        def eig_l(l):
            lam, Q = np.linalg.eigh(self.grid.hamiltonian(l, self.atom.Z, self.v_gs))
            nk = int(np.searchsorted(lam, self.ecut, side="right"))
            return lam[:nk], Q[:, :nk]

        # one dense eigendecomposition per l; LAPACK releases the GIL, so the
        # l run on a thread pool, each with a single BLAS thread
        from concurrent.futures import ThreadPoolExecutor
        workers = max(1, min(16, os.cpu_count() or 1, self.n_l))
        with _blas_threads(1), ThreadPoolExecutor(max_workers=workers) as ex:
            eig = list(ex.map(eig_l, range(self.n_l)))
        self.nkeep = np.array([len(e[0]) for e in eig], dtype=np.int32)
        self.nk_max = _round16(max(1, self.nkeep.max()))
        self.Q = np.zeros((self.n_l, self.N, self.nk_max))
        self.lam = np.zeros((self.n_l, self.nk_max))
        for l, (lam, Q) in enumerate(eig):
            self.Q[l, :, :lam.size] = Q
            self.lam[l, :lam.size] = lam

    def _tables_device(self, device):
        """The same eigenpairs from one batched FP64 eigendecomposition on the
        GPU (H_l formed there from T, r and v = -Z/r + v_gs)."""
        from . import _lib
        lib = _lib.load()
        g = self.grid
        cap = _round16(self.N)
        T = np.ascontiguousarray(g.T, dtype=np.float64)
        v = np.ascontiguousarray(-self.atom.Z / g.r + self.v_gs, dtype=np.float64)
        r = np.ascontiguousarray(g.r, dtype=np.float64)
        Q = np.empty((self.n_l, self.N, cap))
        lam = np.empty((self.n_l, cap))
        nkeep = np.empty(self.n_l, dtype=np.int32)
        need = np.zeros(1, dtype=np.int32)
        rc = lib.ggb_propagator_tables(int(device), self.N, self.n_l, _lib.dptr(r), _lib.dptr(T),
                                       _lib.dptr(v), float(self.ecut), cap, _lib.dptr(Q),
                                       _lib.dptr(lam), _lib.iptr(nkeep), _lib.iptr(need))
        _lib.check(rc, None, "ggb_propagator_tables")
        self.nkeep = nkeep
        self.nk_max = _round16(max(1, int(nkeep.max())))
        self.Q = np.ascontiguousarray(Q[:, :, :self.nk_max])
        self.lam = np.ascontiguousarray(lam[:, :self.nk_max])

    def phase_eigenvalues(self, lam):
        """The eigenvalues whose e^{-i lam dt} the device applies: lam itself, or
        for the Crank-Nicolson forms 2 arg(p(lam dt/2)) / dt."""
        lam = np.asarray(lam, dtype=np.float64)
        if self.propagator == "exp":
            return lam
        n = {"cn": 1, "cn2": 2, "cn3": 3}[self.propagator]
        x = 0.5 * self.dt * lam
        re = np.ones_like(x)
        im = np.zeros_like(x)
        if n >= 1:
            im = im + x
        if n >= 2:
            re = re - 0.5 * x * x
        if n >= 3:
            im = im - x * x * x / 6.0
        return 2.0 * np.arctan2(im, re) / self.dt

    def static_arrays(self):
        """Contiguous arrays in the layouts of ggb_static (include/ggbtd.h)."""
        a = self.angular
        lam = self.phase_eigenvalues(self.lam)
        return dict(r=np.ascontiguousarray(self.grid.r), wr=np.ascontiguousarray(self.grid.wr),
                    v_gs=np.ascontiguousarray(self.v_gs), mask=np.ascontiguousarray(self.mask),
                    Q=np.ascontiguousarray(self.Q), lam=np.ascontiguousarray(lam),
                    nkeep=np.ascontiguousarray(self.nkeep), B=np.ascontiguousarray(a.B),
                    PL=np.ascontiguousarray(a.PL), chi=np.ascontiguousarray(a.x),
                    omega=np.ascontiguousarray(a.omega), m=np.ascontiguousarray(self.m),
                    occ=np.ascontiguousarray(self.occ), **self.reference_orbitals())

    def reference_orbitals(self):
        """Ground-state orbital of each propagated orbital (its starting
        state, up to a global phase): real radial function phi [n_orb, N] in u
        units with unit norm, its l and its
        energy eps = sum_a lam_a |<q_a|phi>|^2 over the retained eigenpairs
        (phi is an eigenvector of its block, so this is its eigenvalue; with
        the Crank-Nicolson propagator the eigenvalue the device's phase uses)."""
        orbs = self.atom.orbitals
        phi = np.zeros((self.n_orb, self.N))
        l0 = np.zeros(self.n_orb, dtype=np.int32)
        eps = np.zeros(self.n_orb)
        for k, (n, l, m, occ, s) in enumerate(orbs):
            u = self.psi0[l, k]
            # u = e^{i theta} v with v real: remove the global phase, unit norm
            th = 0.5 * np.angle(np.sum(u * u))
            v = (u * np.exp(-1j * th)).real
            phi[k] = v / np.linalg.norm(v)
            u = phi[k]
            l0[k] = l
            c = self.Q[l, :, :self.nkeep[l]].T @ u
            eps[k] = float((self.lam[l, :self.nkeep[l]] * c * c).sum() / (c * c).sum())
        return dict(phi=phi, l0=l0, eps=self.phase_eigenvalues(eps))

    def total_acceleration(self, acc, norm, E):
        """Ehrenfest dipole acceleration of all electrons,
        a(t) = -Z sum_k occ_k <cos(theta)/r^2>_k - E(t) sum_k occ_k <psi_k|psi_k>
        (Hartree and LDA forces cancel in the sum over electrons)."""
        return -self.atom.Z * (acc @ self.occ) - E * (norm @ self.occ)

    def table_bytes(self):
        """Bytes of the eigenvector tables streamed per step (FP64)."""
        return int(self.nkeep.sum()) * self.N * 8
