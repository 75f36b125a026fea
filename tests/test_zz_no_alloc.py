"""The allocation-free spectrum entry point (ggb_spectrum_noalloc) against
ggb_spectrum, through the C ABI: two calls in a row, the second with a
smaller input, each bit for bit equal to the allocating entry point (same
kernels, same order of operations); sizes beyond its fixed scratch are
refused with a message and no device work."""
import numpy as np
import pytest

from ggbtddft_amd import _lib

pytestmark = pytest.mark.gpu

def _bind():
    return _lib.load()


def _signal(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) * 0.05
    x = sum(10.0 ** (-0.15 * q) * np.sin(q * 0.057 * t + rng.uniform(0, 6.3)) for q in range(1, 40, 2))
    return np.ascontiguousarray(x + 1e-9 * rng.standard_normal(n))


def _call(fn, x, differentiate, n_pad, orders):
    M = (2 * n_pad + 1) * x.size
    power = np.empty(M)
    inten = np.empty(orders.size)
    rc = fn(0, _lib.dptr(x), x.size, 0.05, int(differentiate), 24.0, n_pad, 0.057,
            _lib.dptr(orders), orders.size, _lib.dptr(power), _lib.dptr(inten))
    _lib.check(rc, None, "spectrum")
    return power, inten


@pytest.mark.parametrize("differentiate", [False, True])
def test_noalloc_matches_allocating_entry_point(differentiate):
    print("case: ggb_spectrum_noalloc vs ggb_spectrum, differentiate=%s" % differentiate, flush=True)
    lib = _bind()
    orders = np.arange(1, 40, 2, dtype=np.float64)
    # two calls in a row, the second smaller: the scratch keeps the first
    # call's contents, and the second must not depend on them
    for n, n_pad in ((27600, 8), (4413, 8)):
        x = _signal(n, n)
        p_ref, i_ref = _call(lib.ggb_spectrum, x, differentiate, n_pad, orders)
        p_na, i_na = _call(lib.ggb_spectrum_noalloc, x, differentiate, n_pad, orders)
        print("n", n, "max |diff| power", np.abs(p_na - p_ref).max(), flush=True)
        assert np.array_equal(p_na, p_ref)
        assert np.array_equal(i_na, i_ref)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    lib = _bind()
    x = _signal(40000, 1)                       # 17 * 40000 > 2^19
    orders = np.arange(1, 5, 2, dtype=np.float64)
    power = np.empty(17 * x.size)
    inten = np.empty(orders.size)
    rc = lib.ggb_spectrum_noalloc(0, _lib.dptr(x), x.size, 0.05, 0, 24.0, 8, 0.057,
                                  _lib.dptr(orders), orders.size, _lib.dptr(power), _lib.dptr(inten))
    assert rc == -1
    assert b"2^19" in lib.ggb_last_error(None)
