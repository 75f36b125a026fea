"""Parity at the configurations' full sizes.

* helium, first configuration, full 4413-step run against the committed
  oracle fixture (tests/golden/he_c1_oracle.npz): dipole acceleration
  relative L2 <= 1e-10, norms within 1e-12;
* argon at the second configuration's size (N=512, Lmax=64, L_H=128): 20
  steps against the oracle, and the size-independent properties (field-free
  stationarity, zero initial potential of the closed shells)."""
import os
import sys

import numpy as np
import pytest

from ggbtddft_amd.pulse import Pulse
from ggbtddft_amd.system import AtomSystem

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_helium_full_run_against_fixture():
    _say("case: He C1 full run vs oracle fixture")
    from ggbtddft_amd.engine import Propagator
    ref = np.load(os.path.join(GOLD, "he_c1_oracle.npz"))
    S = AtomSystem("He", 96, 60.0, 12, 0.05)
    assert list(S.nkeep) == list(ref["nkeep"])
    E = Pulse(1e14, 800.0, 2).midpoints(0.05)
    assert np.abs(E - ref["E_mid"]).max() < 1e-15
    with Propagator(S) as prop:
        prop.set_field(E)
        out = prop.run(0, E.size)
    for key in ("dip", "acc"):
        rel = np.linalg.norm(out[key][:, 0, :] - ref[key]) / np.linalg.norm(ref[key])
        _say(key, "rel L2", rel)
        assert rel < 1e-10, (key, rel)
    dn = np.abs(out["norm"][:, 0, :] - ref["norm"]).max()
    _say("norm max diff", dn)
    assert dn < 1e-12


@pytest.fixture(scope="module")
def argon():
    return AtomSystem("Ar", 512, 200.0, 64, 0.02)


def test_argon_full_size_against_oracle(argon):
    _say("case: Ar C2 size, 20 steps vs oracle")
    from ggbtddft_amd.engine import Propagator
    from oracle import propagate as OP
    O = OP.System("Ar", 512, 200.0, 64, 0.02)
    p = Pulse(2e14, 800.0, 8)
    E = p.midpoints(0.02)[5000:5020]       # inside the pulse, |E| ~ E0/3
    ref = O.run(E)
    with Propagator(argon) as prop:
        prop.set_field(E)
        out = prop.run(0, E.size)
        psi = prop.download_orbitals()
    for key in ("dip", "acc"):
        rel = np.linalg.norm(out[key][:, 0, :] - ref[key]) / np.linalg.norm(ref[key])
        _say(key, "rel L2", rel)
        assert rel < 1e-10, (key, rel)
    assert np.abs(out["norm"][:, 0, :] - ref["norm"]).max() < 1e-12
    assert np.abs(psi - ref["psi"]).max() < 1e-10


def test_argon_field_free_invariants(argon):
    _say("case: Ar C2 size field-free invariants")
    from ggbtddft_amd.engine import Propagator
    with Propagator(argon) as prop:
        dv0 = prop.download_potential()
        prop.set_field(np.zeros(200))
        out = prop.run(0, 200)
        psi = prop.download_orbitals()
    assert np.abs(dv0).max() < 1e-8
    assert np.abs(out["dip"]).max() < 1e-12
    assert np.abs(out["norm"] - out["norm"][0]).max() < 1e-12
    for k in range(argon.n_orb):
        ov = abs(np.vdot(argon.psi0[:, k, :], psi[:, k, :]))
        assert abs(ov - 1.0) < 1e-9, (k, ov)
