"""HHG spectra from the propagator's observables.

``dipole_spectrum`` is the reference's definition (src/scripts/analysis.py:200,
276-278, 447-459): Kaiser window (beta 24, sym=False) on the occupation-
weighted dipole, 8th-order central second difference (internal.py:361-379),
zero padding by n_pad*N_t on each side (N_pad = 8, analysis.py:147),
fftshift(fft(fftshift(.))) and |.|^2.

``acceleration_spectrum`` applies the same window and padding to the Ehrenfest
dipole acceleration (no differencing).  ``harmonic_intensities`` integrates a
spectrum under the super-Gaussian harmonic masks of
argonInSitu/argonInSitu/arg.py:42-67; ``cutoff_energy`` is 3.17 U_p + I_p
(internal.py:203).

``device_spectrum`` computes the same power spectrum and harmonic intensities
on the GPU (``ggb_spectrum``: window, difference, Bluestein DFT of any length
on radix-2 Stockham FFTs, masks), for the serving path's time-to-spectrum.
"""
import threading

import numpy as np
from scipy.signal.windows import kaiser

_FD8 = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)


def second_difference(f, dt):
    f = np.asarray(f, dtype=np.float64)
    out = np.zeros_like(f)
    n = f.size
    if n > 8:
        core = _FD8[0] * f[4:n - 4]
        for s in range(1, 5):
            core = core + _FD8[s] * (f[4 + s:n - 4 + s] + f[4 - s:n - 4 - s])
        out[4:n - 4] = core
    return out / (dt * dt)


def _power(x, n_pad, dt):
    xp = np.pad(x, n_pad * x.size, mode="constant")
    amp = np.fft.fftshift(np.fft.fft(np.fft.fftshift(xp)))
    w = 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(xp.size, dt))
    return w, np.abs(amp) ** 2


def dipole_spectrum(dipole, dt, beta=24.0, n_pad=8):
    d = np.asarray(dipole, dtype=np.float64)
    return _power(second_difference(d * kaiser(d.size, beta, sym=False), dt), n_pad, dt)


def acceleration_spectrum(acc, dt, beta=24.0, n_pad=8):
    a = np.asarray(acc, dtype=np.float64)
    return _power(a * kaiser(a.size, beta, sym=False), n_pad, dt)


def harmonic_intensities(w, S, omega0, orders):
    idx = np.arange(w.size, dtype=np.float64)
    res = []
    for H in np.asarray(orders, dtype=np.float64):
        i0 = np.argmin(np.abs(H * omega0 - w))
        ip = np.argmin(np.abs((H + 0.5) * omega0 - w))
        im = np.argmin(np.abs((H - 0.5) * omega0 - w))
        res.append(np.sum(np.exp(-((idx - i0) / (0.25 * (ip - im))) ** 32) * S))
    return np.array(res)


def cutoff_energy(E0, omega0, ip):
    return 3.17 * (E0 / (2.0 * omega0)) ** 2 + abs(ip)


NOALLOC_MAX_POINTS = 1 << 19   # (2 n_pad + 1) n of ggb_spectrum_noalloc's fixed scratch
NOALLOC_MAX_ORDERS = 256
_noalloc_lock = threading.Lock()  # the fixed scratch is shared by every call on a device


def noalloc_fits(n, n_pad=8, n_orders=0):
    """Whether ggb_spectrum_noalloc takes a series of n samples."""
    return (2 * n_pad + 1) * n <= NOALLOC_MAX_POINTS and n_orders <= NOALLOC_MAX_ORDERS


def device_spectrum(signal, dt, differentiate=False, beta=24.0, n_pad=8, omega0=None,
                    orders=None, device=0, want_power=True, noalloc=False):
    """(w, power, intensities) of ``dipole_spectrum`` (differentiate=True) or
    ``acceleration_spectrum`` computed on the GPU; intensities of ``orders``
    with the masks of ``harmonic_intensities`` (None without orders).
    noalloc: through ggb_spectrum_noalloc (no device allocation, legacy
    stream, same results bit for bit; calls are serialised here)."""
    from . import _lib
    lib = _lib.load()
    x = np.ascontiguousarray(signal, dtype=np.float64)
    n = x.size
    M = (2 * n_pad + 1) * n
    power = np.empty(M) if want_power else None
    ords = None if orders is None else np.ascontiguousarray(orders, dtype=np.float64)
    inten = None if ords is None else np.empty(ords.size)
    args = (int(device), _lib.dptr(x), n, float(dt), int(bool(differentiate)),
            float(beta), int(n_pad), float(omega0 or 0.0),
            None if ords is None else _lib.dptr(ords),
            0 if ords is None else ords.size,
            None if power is None else _lib.dptr(power),
            None if inten is None else _lib.dptr(inten))
    if noalloc:
        with _noalloc_lock:
            rc = lib.ggb_spectrum_noalloc(*args)
        _lib.check(rc, None, "ggb_spectrum_noalloc")
    else:
        rc = lib.ggb_spectrum(*args)
        _lib.check(rc, None, "ggb_spectrum")
    w = 2.0 * np.pi * (np.arange(M) - M // 2) / (M * dt)
    return w, power, inten
