"""Pins the oracle's radial mesh (oracle/mesh.py) against the reference's root
files and closed forms."""
import os

import numpy as np
import pytest

from oracle import mesh as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("N", [32, 64, 128, 256, 512])
def test_roots_match_reference_files(N):
    ref = np.genfromtxt(os.path.join(GOLD, "roots%05d.txt" % N), delimiter=",")
    ref = np.sort(ref[~np.isnan(ref)])
    assert ref.size == N
    np.testing.assert_allclose(M.gl_interior_nodes(N), ref, rtol=0, atol=2e-15)


@pytest.mark.parametrize("N", [33, 96, 200])
def test_roots_are_critical_points_of_legendre(N):
    x = M.gl_interior_nodes(N)
    p, pm1 = M.legendre_p_and_dp(N + 1, x)
    dp = (N + 1) * (x * p - pm1) / (x * x - 1.0)
    assert np.abs(dp).max() < 1e-10 * np.abs(p).max() * N * N


def test_lobatto_weights_and_quadrature():
    N = 40
    m = M.lobatto_mesh(N, 10.0, 1.0)
    x, w = m["x"], m["w"]
    # interior weights = 2 - the two endpoint weights 2/((N+2)(N+1))
    assert abs(w.sum() - (2.0 - 4.0 / ((N + 2) * (N + 1)))) < 1e-13
    # exact for (1-x^2) x^k, k <= 2N - 1 (degree <= 2N+1)
    for k in (0, 3, 10, 2 * N - 1):
        exact = 2.0 / (k + 1) - 2.0 / (k + 3) if k % 2 == 0 else 0.0
        assert abs(np.sum(w * (1 - x * x) * x ** k) - exact) < 1e-13
    # mapped quadrature int_0^R r^2 e^{-r} dr
    m = M.lobatto_mesh(200, 80.0, 1.0)
    assert abs(np.sum(m["wr"] * m["r"] ** 2 * np.exp(-m["r"])) - 2.0) < 1e-12


@pytest.mark.parametrize("Z", [1.0, 2.0])
def test_hydrogenic_levels(Z):
    m = M.lobatto_mesh(256, 100.0 / Z, 1.0)
    for l in (0, 1, 2):
        e = np.linalg.eigvalsh(M.radial_hamiltonian(m, l, Z))
        exact = [-Z * Z / (2.0 * n * n) for n in range(l + 1, l + 4)]
        np.testing.assert_allclose(e[:3], exact, rtol=2e-5)


def test_kinetic_symmetric_and_negative():
    m = M.lobatto_mesh(64, 30.0, 1.0)
    T = m["T"]
    assert np.abs(T - T.T).max() == 0.0
    assert np.linalg.eigvalsh(T).max() < 0.0
