"""The reference-style driver (init file -> output tree) on the CPU with the
oracle-backed stand-in propagator: field convention, batching of in-situ
scans (N_scan perturbed scans and the unperturbed one), the total
acceleration, and the files written -- read back the way the reference's
analysis script reads them (params.cuh macros, per-(orbital, l) channel
files and its channel rule)."""
import os
import re

import numpy as np
import pytest

from fake_backend import OraclePropagator
from ggbtddft_amd import run as R
from ggbtddft_amd.config import RunConfig, read_init
from oracle import propagate as OP

INIT = """atom = %s
N_r = %d
zeta = 1.0
N_l = %d
N_t = %d
dt = 4.0
I0 = 1e14
omega0 = 0.057
USE_HARTREE = 1
IN_SITU = %d
N_scan = 3
"""
DATE = 20261203


def _cfg(atom="He", N=24, nl=4, nt=48, in_situ=0):
    return RunConfig(read_init(INIT % (atom, N, nl, nt, in_situ)))


def _direct(cfg, delay):
    O = OP.System(cfg.atom, cfg.N, cfg.r_max, cfg.lmax, cfg.dt)
    t = cfg.time_grid()
    ref = O.run(cfg.supergaussian_field(t - 0.5 * cfg.dt, delay), channels=True)
    return O, ref, cfg.supergaussian_field(t, delay)


def _parse_params(path):
    """The reference's parseParams: '#define NAME VALUE' lines, values kept
    as text, read back by analysis.py as int/float(value[1:-1])."""
    out = {}
    with open(path) as f:
        for line in f:
            m = re.search("#define +([^ ]+)(?: +(.*))?", line.strip())
            if m:
                out[m.group(1)] = m.group(2)
    return out


# getKSOL / getKSOM of the reference (internal.py:16-42), by target id
KSOL = {1: [0], 3: [0, 1, 1]}
KSOM = {1: [0], 3: [0, 0, 1]}


def _analysis_dipole(base, params, scan):
    """Sum of the channel files the reference's analysis takes for each
    orbital (analysis.py:407-430): l from m (or from l0 when l0 = 0 or
    l0 = m) to N_l - 5, leaving out l = l0; real parts, per orbital."""
    n_kso, n_l, tid = int(params["N_kso"][1:-1]), int(params["N_l"][1:-1]), int(params["targetID"][1:-1])
    n_t = int(params["N_t"][1:-1])
    lext = n_l - 4
    out = np.zeros((n_kso, n_t))
    for k in range(n_kso):
        l0, m0 = KSOL[tid][k], KSOM[tid][k]
        mindx = l0 if (l0 == 0 or l0 == m0) else m0
        for l in sorted(set(range(mindx, lext)) - {l0}):
            f = os.path.join(base, "timeDependent", "observables", "dipoleMoment",
                             "scan%03d" % scan, "kso%03d" % k, "dipL%03d.bin" % l)
            out[k] += np.fromfile(f, dtype=np.complex128).real
    return out


def test_single_scan_run_writes_the_oracle_numbers(tmp_path):
    cfg = _cfg()
    base, scans = R.run(cfg, str(tmp_path / "output"), str(tmp_path / "workingData"),
                        index=3, factory=OraclePropagator, sim_date=DATE)
    assert base == str(tmp_path / "output" / "20261203" / "00003") and len(scans) == 1
    O, ref, e_end = _direct(cfg, None)
    np.testing.assert_allclose(scans[0]["acc"], ref["acc"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(scans[0]["norm"], ref["norm"], rtol=0, atol=1e-13)
    a_ref = -O.Z * (ref["acc"] @ O.occ) - e_end * (ref["norm"] @ O.occ)
    np.testing.assert_allclose(scans[0]["acc_total"], a_ref, rtol=0, atol=1e-13)
    td = os.path.join(base, "timeDependent")
    acc = np.fromfile(os.path.join(td, "observables", "dipoleAcceleration", "scan000", "acc.bin"))
    assert np.array_equal(acc, scans[0]["acc_total"])
    # one file per (orbital, l), holding the channel dipoles; their sum is <z>
    chan = np.stack([np.fromfile(os.path.join(td, "observables", "dipoleMoment", "scan000", "kso000",
                                              "dipL%03d.bin" % l), dtype=np.complex128)
                     for l in range(cfg.lmax + 1)], axis=1)
    np.testing.assert_allclose(chan, ref["chan"][:, 0], rtol=0, atol=1e-13)
    np.testing.assert_allclose(chan.real.sum(axis=1), ref["dip"][:, 0], rtol=0, atol=1e-12)
    E = np.fromfile(os.path.join(td, "field", "E000.bin"))
    assert np.array_equal(E, e_end) and E.size == cfg.n_t
    t = np.fromfile(str(tmp_path / "workingData" / "mesh" / "time.bin"))
    assert np.array_equal(t, cfg.time_grid())
    S = np.fromfile(os.path.join(td, "spectrum", "scan000", "S.bin"))
    assert S.size == 17 * cfg.n_t and np.all(np.isfinite(S))


def test_params_file_reads_back_like_the_reference(tmp_path):
    cfg = _cfg()
    base, _ = R.run(cfg, str(tmp_path / "output"), index=5, factory=OraclePropagator, sim_date=DATE)
    p = _parse_params(os.path.join(base, "params.cuh"))
    # analysis.py:111-143 reads these, as int/float(value[1:-1])
    assert int(p["SIM_DATE"][1:-1]) == DATE and int(p["SIM_INDEX"][1:-1]) == 5
    assert int(p["TD_SIMULATION"][1:-1]) == 1 and int(p["IN_SITU"][1:-1]) == 0
    assert int(p["N_r"][1:-1]) == cfg.N and int(p["N_l"][1:-1]) == cfg.lmax + 1
    assert int(p["N_t"][1:-1]) == cfg.n_t and int(p["N_scan"][1:-1]) == 1
    assert int(p["N_kso"][1:-1]) == 1 and int(p["targetID"][1:-1]) == 1
    assert int(p["N_har"][1:-1]) == 2 * cfg.lmax + 1
    for key, val in (("E0", cfg.E0), ("omega0", cfg.omega0), ("tau0", cfg.tau0),
                     ("R_MASK", cfg.r_mask), ("R_MAX", cfg.r_max), ("dt", cfg.dt)):
        assert abs(float(p[key][1:-1]) - val) <= 1e-14 * abs(val), key
    assert float(p["ETA_MASK"][1:-1]) == 1e-3


def test_analysis_channel_rule_on_the_tree(tmp_path):
    # argon: the analysis reads the reference's three active orbitals (3s, 3p0,
    # 3p1), which the writer puts first; its channel rule applied to the files
    # gives what it gives on the oracle's channels of those orbitals
    cfg = _cfg("Ar", 32, 7, 24)
    base, scans = R.run(cfg, str(tmp_path / "output"), factory=OraclePropagator, sim_date=DATE)
    p = _parse_params(os.path.join(base, "params.cuh"))
    assert int(p["N_kso"][1:-1]) == 3 and int(p["N_kso_ALL"][1:-1]) == 7
    got = _analysis_dipole(base, p, 0)
    O, ref, _ = _direct(cfg, None)
    labels = [(n, l, m) for (n, l, m, _, _) in O.orbs]
    for k, nlm in enumerate([(3, 0, 0), (3, 1, 0), (3, 1, 1)]):
        kk = labels.index(nlm)
        l0, m0 = nlm[1], nlm[2]
        mindx = l0 if (l0 == 0 or l0 == m0) else m0
        want = sum(ref["chan"][:, kk, l].real for l in range(mindx, cfg.lmax + 1 - 4) if l != l0)
        np.testing.assert_allclose(got[k], want, rtol=0, atol=1e-13)
    # every orbital's files together hold its whole dipole
    for kk_file, kk in enumerate(O_order(cfg)):
        tot = sum(np.fromfile(os.path.join(base, "timeDependent", "observables", "dipoleMoment",
                                           "scan000", "kso%03d" % kk_file, "dipL%03d.bin" % l),
                              dtype=np.complex128).real for l in range(cfg.lmax + 1))
        np.testing.assert_allclose(tot, ref["dip"][:, kk], rtol=0, atol=1e-11)


def O_order(cfg):
    from ggbtddft_amd.atoms import Atom
    return Atom(cfg.atom).reference_order()[0]


def test_in_situ_scans_batched_like_alone_and_unperturbed_last(tmp_path):
    cfg = _cfg(in_situ=1)
    base, scans = R.run(cfg, str(tmp_path / "output"), factory=OraclePropagator, max_batch=2,
                        sim_date=DATE)
    # N_scan perturbed scans, then the unperturbed pulse (main.cu:460-467)
    assert len(scans) == 4
    for s, delay in enumerate(list(cfg.delays()) + [None]):
        assert scans[s]["delay"] is None if delay is None else scans[s]["delay"] == delay
        _, ref, e_end = _direct(cfg, delay)
        np.testing.assert_allclose(scans[s]["acc"], ref["acc"], rtol=0, atol=1e-13)
        assert np.array_equal(scans[s]["E"], e_end)
    assert np.array_equal(scans[3]["E"], cfg.supergaussian_field(cfg.time_grid()))
    assert os.path.exists(os.path.join(base, "timeDependent", "field", "E003.bin"))
    p = _parse_params(os.path.join(base, "params.cuh"))
    assert int(p["N_scan"][1:-1]) == 3 and int(p["IN_SITU"][1:-1]) == 1


@pytest.mark.parametrize("atom", ["Ne", "Xe"])
def test_reference_order_puts_active_orbitals_first(atom):
    from ggbtddft_amd.atoms import REFERENCE_ACTIVE, Atom
    A = Atom(atom)
    order, n = A.reference_order()
    assert sorted(order) == list(range(len(A.orbitals)))
    assert [A.orbitals[k][:3] for k in order[:n]] == list(REFERENCE_ACTIVE[atom])
