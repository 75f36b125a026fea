"""Laser pulses (oracle; test infrastructure only).

Unit conversions restate ``init.py:10-14, 255-256``: E0 = sqrt(I / 3.509e16)
for I in W/cm^2; omega0 = 45.5634 / lambda[nm] (hbar c in eV nm over the
Hartree, 1239.84193 / 27.2114).

This propagator's pulse is a sin^2 envelope (the reference uses a super-Gaussian,
``h_initialiseField.cuh:20``):
    E(t) = E0 sin^2(pi t / tau) cos(omega0 (t - tau/2) + cep),  0 <= t <= tau,
    tau  = n_cycles * 2 pi / omega0,
with an optional second colour at harmonic order h (400 nm on 800 nm is h = 2)
and intensity ratio eta: amplitude E0 sqrt(eta), phase phi,
    E2(t) = E0 sqrt(eta) sin^2(pi t / tau) cos(h omega0 (t - tau/2) + phi).
Step n runs t_n = n dt -> t_{n+1}; the laser sub-steps of step n use the
midpoint field E(t_n + dt/2).  n_steps = round(tau / dt).
"""
import numpy as np

INTENSITY_AU = 3.509e16          # init.py:13
NM_HARTREE = 45.5633525          # 1239.84193 eV nm / 27.2113862 eV


def e0_from_intensity(I_wcm2):
    return np.sqrt(I_wcm2 / INTENSITY_AU)


def omega_from_wavelength(nm):
    return NM_HARTREE / nm


def n_steps(n_cycles, omega0, dt):
    return int(round(n_cycles * 2.0 * np.pi / omega0 / dt))


def sin2_field(t, E0, omega0, n_cycles, cep=0.0, second=None):
    """second = (harmonic, intensity_ratio, phase) or None."""
    t = np.asarray(t, dtype=np.float64)
    tau = n_cycles * 2.0 * np.pi / omega0
    inside = (t >= 0.0) & (t <= tau)
    env = np.where(inside, np.sin(np.pi * t / tau) ** 2, 0.0)
    E = E0 * env * np.cos(omega0 * (t - 0.5 * tau) + cep)
    if second is not None:
        h, eta, phi = second
        E = E + E0 * np.sqrt(eta) * env * np.cos(h * omega0 * (t - 0.5 * tau) + phi)
    return E


def midpoint_series(E0, omega0, n_cycles, dt, cep=0.0, second=None, nsteps=None):
    if nsteps is None:
        nsteps = n_steps(n_cycles, omega0, dt)
    t = (np.arange(nsteps) + 0.5) * dt
    return sin2_field(t, E0, omega0, n_cycles, cep, second)
