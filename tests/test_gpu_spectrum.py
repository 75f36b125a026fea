"""Device spectrum (ggb_spectrum) against the oracle's restatement of the
reference's definition (oracle/spectrum.py: analysis.py:447-460,
internal.py:361-379, numpy FFT): power spectra of accelerations and of
dipoles (with the 8th-order second difference), and harmonic-mask
intensities, on lengths that are not powers of two (Bluestein DFT)."""
import numpy as np
import pytest

from ggbtddft_amd import spectrum as spec
from oracle import spectrum as OS

pytestmark = pytest.mark.gpu


def _signal(n, dt, omega0, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) * dt
    x = np.zeros(n)
    for q in range(1, 40, 2):  # odd harmonics, intensities falling over ~12 decades
        x += 10.0 ** (-0.15 * q) * np.sin(q * omega0 * t + rng.uniform(0, 2 * np.pi))
    return x + 1e-9 * rng.standard_normal(n)


@pytest.mark.parametrize("n,n_pad,differentiate", [(1000, 8, False), (4413, 8, False),
                                                    (1001, 3, True), (4096, 8, True)])
def test_power_spectrum_matches_oracle(n, n_pad, differentiate):
    print("case: device spectrum n=%d n_pad=%d differentiate=%s" % (n, n_pad, differentiate),
          flush=True)
    dt, omega0 = 0.05, 0.057
    x = _signal(n, dt, omega0, seed=n)
    if differentiate:
        w_h, p_h = OS.dipole_spectrum(x, dt, n_pad=n_pad)
    else:
        w_h, p_h = OS.acceleration_spectrum(x, dt, n_pad=n_pad)
    w_d, p_d, _ = spec.device_spectrum(x, dt, differentiate=differentiate, n_pad=n_pad)
    assert p_d.shape == p_h.shape
    np.testing.assert_allclose(w_d, w_h, rtol=0, atol=1e-12 * np.abs(w_h).max())
    # both are FP64 FFTs: agreement to rounding of the largest amplitude
    assert np.abs(p_d - p_h).max() <= 1e-11 * p_h.max()


def test_harmonic_intensities_match_oracle():
    print("case: device harmonic intensities n=4413", flush=True)
    n, dt, omega0 = 4413, 0.05, 0.057
    x = _signal(n, dt, omega0, seed=7)
    orders = np.arange(1, 40, 2, dtype=np.float64)
    w_h, p_h = OS.acceleration_spectrum(x, dt)
    ref = OS.harmonic_intensities(w_h, p_h, omega0, orders)
    _, _, got = spec.device_spectrum(x, dt, omega0=omega0, orders=orders, want_power=False)
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-12 * ref.max())
