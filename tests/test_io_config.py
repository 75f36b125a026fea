"""Reference-format inputs/outputs: the init-key reader and the
workingData/output writers."""
import os

import numpy as np
import pytest

from ggbtddft_amd import io
from ggbtddft_amd.config import RunConfig, read_init
from ggbtddft_amd.system import AtomSystem

INIT = """# Target System
atom = Xe
N_r  = 512
zeta = 1.0
N_l = 256
N_t = 16384
dt  = 1.0
I0 = 1.3E14
omega0 = 0.0253
tau0 = 2.0 * np.pi / 0.0253 / 41.3413745758
CEP0 = 0.5 * np.pi
IN_SITU = 1
N_scan = 256
USE_HARTREE = 1
EXP_METHOD = CN
"""


def test_init_reader_matches_reference_conventions():
    c = RunConfig(read_init(INIT))
    assert c.atom == "Xe" and c.N == 512 and c.lmax == 255 and c.n_scan == 256
    assert abs(c.dt - 0.041341374575751) < 1e-15          # 1 as in a.u.
    assert abs(c.E0 - np.sqrt(1.3e14 / 3.509e16)) < 1e-15
    assert abs(c.tau0 - 2 * np.pi / 0.0253) < 1e-9        # fs -> a.u. round trip
    assert abs(c.r_max - 1.5 * c.E0 / 0.0253 ** 2) < 1e-9
    t = c.time_grid()
    assert t[0] == -0.5 * c.n_t * c.dt and abs(t[1] - t[0] - c.dt) < 1e-12
    assert c.propagator == "cn"                            # EXP_METHOD = CN


@pytest.mark.parametrize("value,prop", [("CN", "cn"), ("CN2", "cn2"), ("CN3", "cn3"), ("EIG", "exp"),
                                        ("0", "cn"), ("1", "cn2"), ("2", "cn3"), ("3", "exp"),
                                        ("EXPM", "exp")])
def test_exp_method_maps_to_the_propagator(value, prop):
    # init.py:55-58: CN = 0, CN2 = 1, CN3 = 2, EIG = 3 (default); internal.py:252-318
    c = RunConfig(read_init("atom = He\nEXP_METHOD = %s\n" % value))
    assert c.propagator == prop
    assert RunConfig(read_init("atom = He\n")).propagator == "exp"


def test_init_reader_refuses_code():
    with pytest.raises(ValueError):
        read_init("atom = __import__('os').system('true')")
    with pytest.raises(ValueError):
        read_init("N_r = open('x').read()")


def test_working_data_round_trip(tmp_path):
    S = AtomSystem("Ar", 48, 40.0, 6, 0.05)
    io.write_working_data(str(tmp_path), S, np.arange(10) * 0.05)
    psi = io.read_eigenstates(str(tmp_path), S.atom.target_id, S.n_l, S.n_orb, S.N)
    order, n_active = S.atom.reference_order()
    assert np.array_equal(psi, S.psi0[:, order])
    # the reference's active orbitals (3s, 3p0, 3p1) first, then the core
    occ = np.fromfile(os.path.join(tmp_path, "stateParameters", "occ_active.bin"), dtype=np.int32)
    assert list(occ) == [2, 2, 4, 2, 2, 2, 4] and n_active == 3
    l = np.fromfile(os.path.join(tmp_path, "stateParameters", "l_active.bin"), dtype=np.int32)
    assert list(l[:3]) == [0, 1, 1]
    r = np.fromfile(os.path.join(tmp_path, "mesh", "r.bin"))
    assert np.array_equal(r, S.grid.r)


def test_output_tree(tmp_path):
    S = AtomSystem("He", 32, 30.0, 4, 0.05)
    nt = 20
    chan = np.zeros((nt, 1, S.n_l), dtype=np.complex128)
    chan[:, 0, 1] = np.arange(nt) + 0.5j
    scans = [dict(E=np.ones(nt), chan=chan, acc_total=np.zeros(nt), omega=np.zeros(3),
                  spectrum=np.ones(3))]
    base = io.write_output(str(tmp_path), 7, S, np.arange(nt) * 0.05, scans, sim_date=20261203)
    assert base == os.path.join(str(tmp_path), "20261203", "00007")
    d = np.fromfile(os.path.join(base, "timeDependent", "observables", "dipoleMoment", "scan000",
                                 "kso000", "dipL001.bin"), dtype=np.complex128)
    assert np.array_equal(d, chan[:, 0, 1])
    z = np.fromfile(os.path.join(base, "timeDependent", "observables", "dipoleMoment", "scan000",
                                 "kso000", "dipL003.bin"), dtype=np.complex128)
    assert z.size == nt and not z.any()
    assert os.path.exists(os.path.join(base, "params.cuh"))
    with pytest.raises(ValueError):
        io.write_output(str(tmp_path), 8, S, np.arange(nt) * 0.05, [dict(scans[0], chan=chan[:, :, :2])])
