"""Pins the oracle's ground state (oracle/groundstate.py), propagation
(oracle/propagate.py) and spectra (oracle/spectrum.py)."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import angular as A, atoms, field as F, groundstate as G, mesh as M
from oracle import propagate as P, spectrum as SP

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- ground state
def test_hydrogen_ground_state_exact():
    m = M.lobatto_mesh(96, 60.0, 1.0)
    gs = G.ground_state(m, 1.0, atoms.shells("H"), interacting=False)
    assert abs(gs["eps"][0] + 0.5) < 1e-8
    assert np.abs(gs["v_gs"]).max() == 0.0


def test_helium_lda_self_consistent():
    m = M.lobatto_mesh(384, 60.0, 1.0)
    gs = G.ground_state(m, 2.0, atoms.shells("He"))
    u = gs["u"][0]
    assert abs(np.vdot(u, u).real - 1.0) < 1e-13
    assert abs(gs["rho_u"].sum() - 2.0) < 1e-12
    H = M.radial_hamiltonian(m, 0, 2.0, gs["v_gs"])
    res = H @ u - gs["eps"][0] * u
    assert np.linalg.norm(res) < 1e-6
    # LDA (Slater + PW92) He 1s eigenvalue, complete-basis value -0.5704;
    # the reference's kink quadrature of the Hartree integral converges O(N^-2)
    assert abs(gs["eps"][0] + 0.5704) < 1e-3


def test_argon_levels_and_orthogonality():
    m = M.lobatto_mesh(256, 200.0, 1.0)
    gs = G.ground_state(m, 18.0, atoms.shells("Ar"))
    # NIST LDA reference eigenvalues (Hartree): 1s 2s 2p 3s 3p
    nist = np.array([-113.800134, -10.794172, -8.443439, -0.883384, -0.382330])
    # N = 256: the O(N^-2) Hartree quadrature leaves a few 1e-3 relative
    assert np.abs(gs["eps"] / nist - 1).max() < 5e-3
    u = gs["u"]
    assert abs(np.vdot(u[0], u[1])) < 1e-12 and abs(np.vdot(u[1], u[3])) < 1e-12


def test_frozen_density_fixture_u_units():
    # densityFrozen.bin holds density_u / (4 pi) of the Xe 1s..3d core
    d = np.fromfile(os.path.join(GOLD, "densityFrozen.bin"), dtype=np.float64)
    assert d.size == 512
    assert abs(4.0 * np.pi * d.sum() - 28.0) < 1e-9
    assert d.min() > -1e-17       # round-off of the reference's subtraction


# ---------------------------------------------------------------- propagation
@pytest.fixture(scope="module")
def he_small():
    return P.System("He", 48, 40.0, 6, 0.05)


def test_field_free_is_stationary(he_small):
    S = he_small
    res = S.run(np.zeros(40))
    assert np.abs(res["dip"]).max() < 1e-13
    assert np.abs(res["norm"] - 1.0).max() < 1e-12
    overlap = abs(np.vdot(S.psi0, res["psi"]))
    assert abs(overlap - 1.0) < 1e-10


def test_initial_potential_vanishes_for_closed_shells():
    S = P.System("Ar", 64, 40.0, 6, 0.05)
    dv = S.potential(S.theta_state(S.psi0))
    assert np.abs(dv).max() < 1e-9


def test_kinetic_is_matrix_exponential():
    S = P.System("He", 24, 20.0, 3, 0.05, ecut_factor=1e9)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=S.psi0.shape) + 1j * rng.normal(size=S.psi0.shape)
    ref = psi.copy()
    S.kinetic(psi)
    for l in range(S.nl):
        H = M.radial_hamiltonian(S.mesh, l, S.Z, S.v_gs)
        ref[l] = (sl.expm(-1j * S.dt * H) @ ref[l].T).T
    assert np.abs(psi - ref).max() < 1e-12


def test_laser_rotations_converge_to_exponential():
    # Le(h) Lo(h) Lo(h) Le(h) = expm(-i 2a C) + O(a^3)
    S = P.System("H", 16, 10.0, 8, 0.05, interacting=False)
    rng = np.random.default_rng(2)
    i = 12
    C = np.diag(A.c_lm(np.arange(8), 0), 1)
    C = C + C.T
    errs = []
    x0 = rng.normal(size=9) + 1j * rng.normal(size=9)
    for E in (0.4, 0.2, 0.1):
        psi = np.zeros_like(S.psi0)
        psi[:, 0, i] = x0
        for par in (0, 1, 1, 0):
            S.laser(psi, E, par)
        a = 0.5 * S.dt * E * S.r[i]
        errs.append(np.abs(psi[:, 0, i] - sl.expm(-1j * 2 * a * C) @ x0).max())
    assert 7.0 < errs[0] / errs[1] < 9.0 and 7.0 < errs[1] / errs[2] < 9.0


def test_norm_conserved_without_absorber():
    S = P.System("Ne", 48, 30.0, 6, 0.05, cap=False)
    w = F.omega_from_wavelength(800.0)
    Em = F.midpoint_series(F.e0_from_intensity(3e14), w, 1, 0.05)[:80]
    res = S.run(Em)
    assert np.abs(res["norm"] - 1.0).max() < 1e-10


def test_hydrogen_static_polarisability():
    # adiabatic switch-on of a weak static field: <z> -> -alpha E, alpha = 9/2
    S = P.System("H", 64, 60.0, 6, 0.1, interacting=False, cap=False)
    Es, T = 1e-3, 120.0
    n = int(2 * T / S.dt)
    t = (np.arange(n) + 0.5) * S.dt
    E = Es * np.where(t < T, np.sin(0.5 * np.pi * t / T) ** 2, 1.0)
    res = S.run(E)
    d = res["dip"][n // 2:, 0]
    alpha = -d.mean() / Es
    assert abs(alpha - 4.5) < 0.05


# ---------------------------------------------------------------- spectra
def test_fd_second_derivative():
    dt = 0.05
    t = np.arange(400) * dt
    f = np.sin(0.7 * t)
    d2 = SP.d2fdx2(f, dt)
    assert np.abs(d2[4:-4] + 0.49 * f[4:-4]).max() < 1e-9


def test_pure_tone_spectrum_peak():
    dt, n, w0 = 0.05, 4096, 1.3
    a = np.cos(w0 * np.arange(n) * dt)
    w, S = SP.acceleration_spectrum(a, dt)
    pos = w > 0
    assert abs(w[pos][np.argmax(S[pos])] - w0) < 2 * np.pi / (17 * n * dt)
    wd, Sd = SP.dipole_spectrum(a, dt)
    # the d^2/dt^2 weighting (omega^2) tilts the Kaiser main lobe by ~2 bins
    assert abs(wd[wd > 0][np.argmax(Sd[wd > 0])] - w0) < 3 * 2 * np.pi / (17 * n * dt)


def test_harmonic_masks_pick_odd_harmonics():
    dt, w0 = 0.05, 0.057
    t = np.arange(65536) * dt    # long enough that the beta-24 lobe < w0
    a = np.cos(3 * w0 * t) + 0.1 * np.cos(5 * w0 * t)
    w, S = SP.acceleration_spectrum(a, dt)
    I = SP.harmonic_intensities(w, S, w0, [3, 4, 5])
    assert I[0] > 50 * I[1] and I[2] > 50 * I[1]
    assert 0.005 < I[2] / I[0] < 0.02
