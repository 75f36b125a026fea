"""Pins the oracle's angular tables (oracle/angular.py): scipy's associated
Legendre functions, orthonormality, the cos(theta) couplings and the
reference's Clebsch-Gordan tables."""
import os
from math import factorial

import numpy as np
import pytest
import scipy.special as sp

from oracle import angular as A

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("m", [0, 1, 2, 3])
def test_theta_lm_matches_scipy(m):
    x = np.linspace(-0.995, 0.995, 23)
    T = A.assoc_legendre_normalised(20, m, x)
    for l in range(m, 21):
        norm = np.sqrt((2 * l + 1) / 2.0 * factorial(l - m) / factorial(l + m))
        ref = (-1) ** m * norm * sp.lpmv(m, l, x)
        np.testing.assert_allclose(T[:, l], ref, rtol=1e-11, atol=1e-12)
    assert np.all(T[:, :m] == 0.0)


@pytest.mark.parametrize("lmax,m", [(12, 0), (12, 1), (64, 0), (64, 2), (96, 1)])
def test_transform_orthonormal(lmax, m):
    # tables built in extended precision and rounded once: B^T B = I to a few
    # ulp and no systematic bias on the diagonal (a double-precision build has
    # ~3e-13 at Lmax = 64, which drifts the norm step after step)
    chi, om = A.theta_grid(2 * lmax + 1, extended=True)
    B = A.transform_matrix(lmax, m, chi, om)
    G = (B.T @ B)[m:, m:] - np.eye(lmax + 1 - m)
    assert np.abs(G).max() < 2e-15
    assert abs(np.diag(G).mean()) < 2e-16


@pytest.mark.parametrize("m", [0, 1, 2])
def test_cos_theta_is_tridiagonal_c_lm(m):
    lmax = 16
    chi, om = A.theta_grid(lmax + 2)
    B = A.transform_matrix(lmax, m, chi, om)
    Cq = B.T @ (chi[:, None] * B)
    l = np.arange(m, lmax)
    np.testing.assert_allclose(np.diag(Cq, 1)[m:], A.c_lm(l, m), atol=1e-13)
    Cq[np.arange(lmax), np.arange(1, lmax + 1)] = 0.0
    Cq[np.arange(1, lmax + 1), np.arange(lmax)] = 0.0
    assert np.abs(Cq).max() < 1e-13


def _cg_rows(name):
    rows = []
    for line in open(os.path.join(GOLD, name)):
        parts = [p for p in line.replace(",", " ").split() if p]
        if len(parts) == 4:
            rows.append(tuple(float(p) for p in parts))
    return rows


@pytest.mark.parametrize("name", ["lo000.txt", "lo002.txt"])
def test_multipole_matrix_against_reference_cg_tables(name):
    # <l1 0|P_L|l2 0> = sqrt((2l1+1)(2l2+1)) <l1 0 l2 0|L 0>^2 / (2L+1)
    lmax = 8
    chi, om = A.theta_grid(2 * lmax + 1)
    B = A.transform_matrix(lmax, 0, chi, om)
    PL = A.legendre_table(2 * lmax, chi)
    rows = _cg_rows(name)
    assert len(rows) >= 8
    for l1, l2, L, cg in rows:
        l1, l2, L = int(l1), int(l2), int(L)
        got = B[:, l1] @ (PL[:, L] * B[:, l2])
        exp = np.sqrt((2 * l1 + 1) * (2 * l2 + 1)) * cg * cg / (2 * L + 1)
        assert abs(got - exp) < 1e-13, (l1, l2, L, got, exp)


def test_pair_lists_cover_tridiagonal():
    for m in (0, 1, 2):
        even, odd = A.pair_lists(10, m)
        assert sorted(even + odd) == list(range(m, 10))
        assert all((l - m) % 2 == 0 for l in even)
