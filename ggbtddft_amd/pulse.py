"""Laser fields.

Conversions follow init.py:255-256 (E0 = sqrt(I / 3.509e16)) and
omega0 = 45.5634 / lambda[nm].  The envelope is sin^2 over n_cycles optical
periods, centred at tau/2:
    E(t) = E0 sin^2(pi t/tau) cos(omega0 (t - tau/2) + cep)
plus an optional second colour (order h, intensity ratio eta, phase phi) of
amplitude E0 sqrt(eta) under the same envelope (the in-situ perturbation of
h_initialiseField.cuh:17-23, here at 2 omega0 for 400 nm on 800 nm).
Step n of the propagator uses E at its midpoint (n + 1/2) dt.
"""
import math

import numpy as np

from .atoms import INTENSITY_AU, NM_TO_HARTREE


class Pulse:
    def __init__(self, intensity_wcm2, wavelength_nm, n_cycles, cep=0.0, second=None):
        self.intensity = float(intensity_wcm2)
        self.wavelength = float(wavelength_nm)
        self.n_cycles = float(n_cycles)
        self.cep = float(cep)
        self.second = second  # (harmonic order, intensity ratio, phase) or None
        self.E0 = math.sqrt(self.intensity / INTENSITY_AU)
        self.omega0 = NM_TO_HARTREE / self.wavelength
        self.tau = self.n_cycles * 2.0 * math.pi / self.omega0

    def n_steps(self, dt):
        return int(round(self.tau / dt))

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        env = np.where((t >= 0.0) & (t <= self.tau), np.sin(math.pi * t / self.tau) ** 2, 0.0)
        ph = self.omega0 * (t - 0.5 * self.tau)
        E = self.E0 * env * np.cos(ph + self.cep)
        if self.second is not None:
            h, eta, phi = self.second
            E = E + self.E0 * math.sqrt(eta) * env * np.cos(h * ph + phi)
        return E

    def midpoints(self, dt, nsteps=None):
        n = self.n_steps(dt) if nsteps is None else int(nsteps)
        return self((np.arange(n) + 0.5) * dt)

    def ponderomotive(self):
        return self.E0 ** 2 / (4.0 * self.omega0 ** 2)
