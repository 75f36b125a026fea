"""HHG spectra (oracle; test infrastructure only).

``dipole_spectrum`` restates the reference's single-run spectrum
(``analysis.py:200, 276-278, 447-459``): Kaiser window (beta 24, sym=False),
8th-order central second difference (``internal.py:361-379``), zero padding by
n_pad * N_t on each side (N_pad = 8, ``analysis.py:147``), then
fftshift(fft(fftshift(.))) and |.|^2.

``acceleration_spectrum`` is this project's form from the Ehrenfest
acceleration: same window and padding, no differentiation.

``harmonic_intensities`` restates the harmonic masks of
``argonInSitu/argonInSitu/arg.py:42-67``: for order q the mask
exp(-((i - i_q)/(0.25 (i_{q+1/2} - i_{q-1/2})))^32) on the frequency index,
summed against the spectrum.

Pinned by: a pure tone (peak at its frequency), Parseval for the padded FFT,
and the FD stencil on polynomials/sinusoids (library routines
scipy.signal.windows.kaiser and numpy.fft are present in the image).
"""
import numpy as np
from scipy.signal.windows import kaiser


def d2fdx2(f, dx):
    """8th-order second derivative, zero in the first/last four samples."""
    c = (-205.0 / 72.0, 8.0 / 5.0, -0.2, 8.0 / 315.0, -1.0 / 560.0)
    f = np.asarray(f, dtype=np.float64)
    out = np.zeros_like(f)
    n = f.size
    k = np.arange(4, n - 4)
    acc = c[0] * f[k]
    for s in range(1, 5):
        acc = acc + c[s] * (f[k + s] + f[k - s])
    out[k] = acc
    return out / dx ** 2


def frequency_axis(n_total, dt):
    return 2.0 * np.pi * np.fft.fftshift(np.fft.fftfreq(n_total, dt))


def _padded_power(x, n_pad):
    n = x.size
    xp = np.pad(x, n_pad * n, mode="constant")
    s = np.fft.fftshift(np.fft.fft(np.fft.fftshift(xp)))
    return s


def dipole_spectrum(D, dt, beta=24.0, n_pad=8):
    D = np.asarray(D, dtype=np.float64)
    win = kaiser(D.size, beta=beta, sym=False)
    s = _padded_power(d2fdx2(D * win, dt), n_pad)
    w = frequency_axis(s.size, dt)
    return w, np.abs(s) ** 2


def acceleration_spectrum(a, dt, beta=24.0, n_pad=8):
    a = np.asarray(a, dtype=np.float64)
    win = kaiser(a.size, beta=beta, sym=False)
    s = _padded_power(a * win, n_pad)
    w = frequency_axis(s.size, dt)
    return w, np.abs(s) ** 2


def harmonic_intensities(w, S, omega0, orders):
    orders = np.asarray(orders, dtype=np.float64)
    idx = np.arange(w.size, dtype=np.float64)
    out = np.empty(orders.size)
    for q, H in enumerate(orders):
        i0 = np.argmin(np.abs(H * omega0 - w))
        ip = np.argmin(np.abs((H + 0.5) * omega0 - w))
        im = np.argmin(np.abs((H - 0.5) * omega0 - w))
        mask = np.exp(-((idx - i0) / (0.25 * (ip - im))) ** 32)
        out[q] = np.sum(mask * S)
    return out


def cutoff_energy(E0, omega0, ip):
    """3.17 U_p + I_p with U_p = E0^2/(4 omega0^2) (internal.py:203)."""
    return 3.17 * (E0 / (2.0 * omega0)) ** 2 + abs(ip)
