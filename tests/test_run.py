"""The reference-style driver (init file -> output tree) on the CPU with the
oracle-backed stand-in propagator: field convention, batching of in-situ
scans, the total acceleration and the files written."""
import os

import numpy as np

from fake_backend import OraclePropagator
from ggbtddft_amd import run as R
from ggbtddft_amd.config import RunConfig, read_init
from oracle import propagate as OP

INIT = """atom = He
N_r = 24
zeta = 1.0
N_l = 4
N_t = 48
dt = 4.0
I0 = 1e14
omega0 = 0.057
USE_HARTREE = 1
IN_SITU = %d
N_scan = 3
"""


def _direct(cfg, delay):
    O = OP.System("He", cfg.N, cfg.r_max, cfg.lmax, cfg.dt)
    t = cfg.time_grid()
    return O, O.run(cfg.supergaussian_field(t - 0.5 * cfg.dt, delay)), cfg.supergaussian_field(t, delay)


def test_single_scan_run_writes_the_oracle_numbers(tmp_path):
    cfg = RunConfig(read_init(INIT % 0))
    base, scans = R.run(cfg, str(tmp_path / "output"), str(tmp_path / "workingData"),
                        index=3, factory=OraclePropagator)
    assert base.endswith("00003") and len(scans) == 1
    O, ref, e_end = _direct(cfg, None)
    np.testing.assert_allclose(scans[0]["acc"], ref["acc"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(scans[0]["norm"], ref["norm"], rtol=0, atol=1e-13)
    a_ref = -O.Z * (ref["acc"] @ O.occ) - e_end * (ref["norm"] @ O.occ)
    np.testing.assert_allclose(scans[0]["acc_total"], a_ref, rtol=0, atol=1e-13)
    td = os.path.join(base, "timeDependent")
    acc = np.fromfile(os.path.join(td, "observables", "dipoleAcceleration", "scan000", "acc.bin"))
    assert np.array_equal(acc, scans[0]["acc_total"])
    d = np.fromfile(os.path.join(td, "observables", "dipoleMoment", "scan000", "kso000",
                                 "dipL000.bin"), dtype=np.complex128)
    np.testing.assert_allclose(d.real, ref["dip"][:, 0], rtol=0, atol=1e-13)
    E = np.fromfile(os.path.join(td, "field", "E000.bin"))
    assert np.array_equal(E, e_end) and E.size == cfg.n_t
    t = np.fromfile(str(tmp_path / "workingData" / "mesh" / "time.bin"))
    assert np.array_equal(t, cfg.time_grid())
    S = np.fromfile(os.path.join(td, "spectrum", "scan000", "S.bin"))
    assert S.size == 17 * cfg.n_t and np.all(np.isfinite(S))


def test_in_situ_scans_batched_like_alone(tmp_path):
    cfg = RunConfig(read_init(INIT % 1))
    _, scans = R.run(cfg, str(tmp_path / "output"), factory=OraclePropagator, max_batch=2)
    assert len(scans) == 3
    for s, delay in enumerate(cfg.delays()):
        _, ref, _ = _direct(cfg, delay)
        np.testing.assert_allclose(scans[s]["acc"], ref["acc"], rtol=0, atol=1e-13)
    assert os.path.exists(tmp_path / "output" / "00000" / "timeDependent" / "field" / "E002.bin")
