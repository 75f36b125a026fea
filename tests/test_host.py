"""CPU tests of the host side: the product precompute against the oracle, and
the C ABI library (loads, exports every declared symbol, rejects bad input
before touching a device)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ggbtddft_amd import _lib
from ggbtddft_amd.atoms import Atom
from ggbtddft_amd.pulse import Pulse
from ggbtddft_amd.system import AtomSystem
from oracle import field as OF
from oracle import propagate as OP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from ggbtddft_amd import build
        build.build()
    return _lib.load()


def test_header_symbols_exported(lib):
    hdr = open(os.path.join(ROOT, "include", "ggbtd.h")).read()
    names = set(re.findall(r"\b(ggb_[a-z_0-9]+)\s*\(", hdr))
    assert names == set(_lib.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n
    assert lib.ggb_abi_version() == _lib.GGB_ABI_VERSION


def test_create_rejects_bad_sizes_without_device(lib):
    p = _lib.Problem()
    p.abi_version = _lib.GGB_ABI_VERSION
    p.n_r, p.n_l, p.n_theta, p.n_har, p.n_orb, p.n_inst, p.n_m, p.nk_max = 64, 9, 17, 17, 1, 1, 1, 17
    p.dt = 0.05
    ctx = ctypes.c_void_p()
    rc = lib.ggb_create(ctypes.byref(p), ctypes.byref(ctx))
    assert rc == -1 and not ctx.value                 # nk_max not a multiple of 16
    assert b"out of range" in lib.ggb_last_error(None)
    p.nk_max = 64
    p.abi_version = 99
    assert lib.ggb_create(ctypes.byref(p), ctypes.byref(ctx)) == -1
    assert lib.ggb_create(None, ctypes.byref(ctx)) == -1


def test_null_context_calls_fail_cleanly(lib):
    assert lib.ggb_run(None, 0, 1, None, None, None) == -1
    assert lib.ggb_upload_orbitals(None, None) == -1
    lib.ggb_destroy(None)


@pytest.mark.parametrize("atom,N,R,lmax,dt", [("He", 96, 60.0, 12, 0.05), ("Ar", 96, 60.0, 8, 0.02),
                                              ("Kr", 96, 60.0, 6, 0.02)])
def test_precompute_matches_oracle(atom, N, R, lmax, dt):
    S = AtomSystem(atom, N, R, lmax, dt)
    O = OP.System(atom, N, R, lmax, dt)
    assert np.abs(S.grid.r - O.r).max() < 1e-13 * R
    assert np.abs(S.grid.T - O.mesh["T"]).max() <= 1e-12 * np.abs(O.mesh["T"]).max()
    assert np.abs(S.gs.eps - O.gs["eps"]).max() < 1e-10 * max(1.0, np.abs(O.gs["eps"]).max())
    assert np.abs(S.v_gs - O.v_gs).max() < 1e-10
    assert np.abs(S.psi0 - O.psi0).max() < 1e-10
    assert list(S.nkeep) == O.nkeep
    for l in range(S.n_l):
        nk = S.nkeep[l]
        assert np.abs(S.lam[l, :nk] - O.lam[l]).max() < 1e-9
        # eigenvectors up to sign
        sgn = np.sign(np.sum(S.Q[l, :, :nk] * O.Q[l], axis=0))
        assert np.abs(S.Q[l, :, :nk] * sgn - O.Q[l]).max() < 1e-8
        assert np.all(S.Q[l, :, nk:] == 0.0)
    for m, B in O.B.items():
        assert np.abs(S.angular.B[m] - B).max() < 1e-13
    assert np.abs(S.angular.PL - O.PL).max() < 1e-13
    assert np.abs(S.mask - O.mask).max() < 1e-15


def test_orbital_sets():
    ar = Atom("argon")
    assert ar.labels == ["1s0", "2s0", "2p0", "3s0", "3p0", "2p1", "3p1"]
    assert [o[3] for o in ar.orbitals] == [2, 2, 2, 2, 2, 4, 4]
    kr = Atom("Kr")
    assert len(kr.orbitals) == 13 and kr.n_electrons == 36
    assert sum(o[3] for o in kr.orbitals) == 36


def test_pulse_matches_oracle_and_configs():
    p = Pulse(1e14, 800.0, 2)
    assert p.n_steps(0.05) == 4413
    assert Pulse(2e14, 800.0, 8).n_steps(0.02) == 44128
    assert abs(p.E0 - 0.0533836563195185) < 1e-12
    Eo = OF.midpoint_series(p.E0, p.omega0, 2, 0.05)
    assert np.abs(p.midpoints(0.05) - Eo).max() < 1e-15
    q = Pulse(1.5e14, 800.0, 8, second=(2, 1e-3, 0.7))
    Eq = OF.midpoint_series(q.E0, q.omega0, 8, 0.02, second=(2, 1e-3, 0.7))
    assert np.abs(q.midpoints(0.02) - Eq).max() < 1e-15


@pytest.mark.parametrize("atom,N,R,lmax", [("He", 48, 30.0, 4), ("Ne", 48, 30.0, 4)])
def test_reference_orbitals_are_unit_eigenvectors(atom, N, R, lmax):
    # the FP32 mixed form advances each orbital's ground-state component
    # exactly: phi must be the starting orbital up to a global phase, of unit
    # norm, and an eigenvector of the retained spectral propagator
    S = AtomSystem(atom, N, R, lmax, 0.05)
    ref = S.reference_orbitals()
    for k, (n, l, m, occ, s) in enumerate(S.atom.orbitals):
        phi = ref["phi"][k]
        assert ref["l0"][k] == l
        assert abs(np.linalg.norm(phi) - 1.0) < 1e-14
        assert abs(abs(np.vdot(phi, S.psi0[l, k])) - 1.0) < 1e-12
        nk = S.nkeep[l]
        c = S.Q[l, :, :nk].T @ phi
        assert abs(c @ c - 1.0) < 1e-10          # inside the retained space
        H = S.grid.hamiltonian(l, S.atom.Z, S.v_gs)
        assert np.linalg.norm(H @ phi - ref["eps"][k] * phi) < 1e-6 * abs(ref["eps"][k])


def test_spectrum_rejects_bad_arguments_without_device(lib):
    # argument checks come before any device call
    x = np.zeros(8)
    rc = lib.ggb_spectrum(0, _lib.dptr(x), 8, 0.05, 0, 24.0, 8, 0.0, None, 0, None, None)
    assert rc == -1  # GGB_E_ARG
    assert b"ggb_spectrum" in lib.ggb_last_error(None)


def test_ground_state_rejects_bad_arguments_without_device(lib):
    # argument checks come before any device call
    N = 8
    z = np.zeros(N)
    zz = np.zeros((N, N))
    sl = np.zeros(1, dtype=np.int32)
    it = np.zeros(1, dtype=np.int32)
    conv = np.zeros(1, dtype=np.int32)
    occ = np.ones(1)
    out = np.zeros(N)
    base = dict(n_r=N, n_shell=1, sweeps=8, max_iter=10, interacting=1, device=0, Z=2.0,
                mix=0.3, gamma=2.0, tol=1e-11)
    for bad in (dict(n_r=2), dict(n_r=4096), dict(n_shell=0), dict(sweeps=0), dict(mix=1.5),
                dict(gamma=0.0), dict(tol=0.0), dict(Z=-1.0)):
        p = _lib.GsProblem(**{**base, **bad})
        rc = lib.ggb_ground_state(ctypes.byref(p), _lib.iptr(sl), _lib.dptr(occ), _lib.dptr(z),
                                  _lib.dptr(z), _lib.dptr(zz), _lib.dptr(zz), _lib.dptr(z),
                                  _lib.dptr(out), _lib.dptr(occ.copy()), _lib.dptr(out.copy()),
                                  _lib.dptr(out.copy()), _lib.iptr(it), _lib.iptr(conv))
        assert rc == -1, bad
        assert b"ggb_ground_state" in lib.ggb_last_error(None)
    # nine shells of one l exceed the per-l limit
    p = _lib.GsProblem(**{**base, "n_shell": 9})
    sl9 = np.zeros(9, dtype=np.int32)
    occ9 = np.ones(9)
    rc = lib.ggb_ground_state(ctypes.byref(p), _lib.iptr(sl9), _lib.dptr(occ9), _lib.dptr(z),
                              _lib.dptr(z), _lib.dptr(zz), _lib.dptr(zz), _lib.dptr(z),
                              _lib.dptr(np.zeros(9 * N)), _lib.dptr(np.zeros(9)), _lib.dptr(out.copy()),
                              _lib.dptr(out.copy()), _lib.iptr(it), _lib.iptr(conv))
    assert rc == -1
