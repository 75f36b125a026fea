"""Batched scans of uneven instances (different grids, intensities,
wavelengths and pulse lengths).

The reference runs the scans of one grid one after another
(main.cu:460-481).  Here a scan is a list of requests; instances of one static
key (atom, N, r_max, Lmax, dt) share a context's slots and tables, and the
instances of different keys run side by side (see ``serving`` and
``ScanGroup``).

``scan_requests`` draws the batched-scan workload: n instances from a seed
over He and Ne (uniform), intensity log-uniform in [5e13, 5e14] W/cm^2,
wavelength uniform in [600, 2000] nm, sin^2 length uniform integer in
[4, 16] cycles (so step counts are uneven), and the grid from the quiver
radius alpha = E0 / omega0^2: N = 256, Lmax = 32, r_max = 100 below 25 a.u.,
N = 384, Lmax = 48, r_max = 150 below 60 a.u., else N = 512, Lmax = 64,
r_max = 200; dt = 0.05.
"""
import numpy as np

from .pulse import Pulse
from .serving import Request

GRIDS = ((25.0, 256, 32, 100.0), (60.0, 384, 48, 150.0), (np.inf, 512, 64, 200.0))


def grid_for(pulse):
    """(N, Lmax, r_max) of the quiver radius alpha = E0 / omega0^2."""
    alpha = pulse.E0 / pulse.omega0 ** 2
    for lim, nr, lmax, rmax in GRIDS:
        if alpha < lim:
            return nr, lmax, rmax
    raise AssertionError


def scan_requests(n, seed=0, dt=0.05, cycle_scale=1.0):
    """The draw above; cycle_scale < 1 shortens every pulse by that factor
    (same grids and fields' shapes, proportionally uneven step counts) for a
    bounded timing run."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        atom = "He" if rng.uniform() < 0.5 else "Ne"
        I = float(np.exp(rng.uniform(np.log(5e13), np.log(5e14))))
        lam = float(rng.uniform(600.0, 2000.0))
        cyc = int(rng.integers(4, 17))
        nr, lmax, rmax = grid_for(Pulse(I, lam, cyc))
        reqs.append(Request(rid=k, atom=atom, intensity=I, wavelength=lam,
                            n_cycles=cyc * float(cycle_scale), N=nr, r_max=rmax, lmax=lmax, dt=dt))
    return reqs
