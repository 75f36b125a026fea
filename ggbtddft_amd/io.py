"""Reference-compatible files (raw little-endian, numpy ``tofile`` order).

Writers follow the layouts the reference's loaders and analysis read:

workingData/mesh/{r,P,drdx}.bin                f64 x N       (init.py:356-358 -> h_initialiseMesh.cuh)
workingData/mesh/time.bin                      f64 x N_t     (init.py:359)
workingData/eigenStates/active/TTT/eigenStatesLlll.bin
                                               c128 x (n_orb*N), [orb][r] per l (init.py:373-375)
workingData/eigenStates/active/TTT/eigenValues.bin  c128 x n_orb (init.py:371)
workingData/stateParameters/{l,m,occ}_active.bin    int32 x n_orb (init.py:377-379)
workingData/stateParameters/densityFrozen.bin  f64 x N       (init.py:380; zeros: all-electron)

output/DDDDDDDD/IIIII/params.cuh                              (init.py:493-608, writeParametersFile)
output/DDDDDDDD/IIIII/static/mesh/{r,drdx,time}.bin           (h_writeMesh.cuh:1-18)
output/DDDDDDDD/IIIII/timeDependent/field/ESSS.bin   f64 x N_t  (h_writeDipoleMoment.cuh:73-87)
output/DDDDDDDD/IIIII/timeDependent/observables/dipoleMoment/scanSSS/ksoKKK/dipLlll.bin
                                               c128 x N_t    (h_writeDipoleMoment.cuh:12-22)

DDDDDDDD is the run date (SIM_DATE, init.py:385) and IIIII the run index.
The dipole files hold the per-(orbital, l) channel dipoles
<x_l| r (c(l,m) x_{l+1} + c(l-1,m) x_{l-1})> computed on the device
(ggb_set_channels), as the reference's dipoleMoment.cuh:1-47 does.

Orbital order: the propagator runs every orbital of the atom (all electrons);
the reference runs a few valence orbitals over a frozen core and its
analysis script indexes its files by that set (getKSOL/getKSOM,
internal.py:16-42; occupations hard-wired at analysis.py:203-204).  The
writers therefore put the reference's active orbitals first, in its order
(kso000.., the workingData orbital lists likewise), then the core orbitals;
params.cuh's N_kso counts the reference's active orbitals, so its analysis
reads exactly those, and N_kso_ALL counts every orbital file written.

Beside them: the Ehrenfest acceleration and the spectrum as
dipoleAcceleration/scanSSS/acc.bin (f64) and spectrum/scanSSS/{w,S}.bin (f64).
"""
import datetime
import os

import numpy as np


def _w(path, arr, dtype):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(arr, dtype=dtype).tofile(path)


def sim_date_today():
    """YYYYMMDD as the reference forms SIM_DATE (init.py:385)."""
    d = datetime.date.today()
    return int("%04d%02d%02d" % (d.year, d.month, d.day))


def write_working_data(root, system, time_grid):
    g = system.grid
    order, _ = system.atom.reference_order()
    orbs = [system.atom.orbitals[k] for k in order]
    _w(os.path.join(root, "mesh", "r.bin"), g.r, np.float64)
    _w(os.path.join(root, "mesh", "P.bin"), g.P, np.float64)
    _w(os.path.join(root, "mesh", "drdx.bin"), g.drdx, np.float64)
    _w(os.path.join(root, "mesh", "time.bin"), time_grid, np.float64)
    tid = "%03d" % system.atom.target_id
    for l in range(system.n_l):
        _w(os.path.join(root, "eigenStates", "active", tid, "eigenStatesL%03d.bin" % l),
           system.psi0[l][order], np.complex128)
    eps = np.array([system.gs.eps[o[4]] for o in orbs], dtype=np.complex128)
    _w(os.path.join(root, "eigenStates", "active", tid, "eigenValues.bin"), eps, np.complex128)
    sp = os.path.join(root, "stateParameters")
    _w(os.path.join(sp, "l_active.bin"), [o[1] for o in orbs], np.int32)
    _w(os.path.join(sp, "m_active.bin"), [o[2] for o in orbs], np.int32)
    _w(os.path.join(sp, "occ_active.bin"), [int(o[3]) for o in orbs], np.int32)
    _w(os.path.join(sp, "densityFrozen.bin"), np.zeros(system.N), np.float64)


def read_eigenstates(root, target_id, n_l, n_orb, N):
    """Load workingData/eigenStates/active/TTT/eigenStatesLlll.bin into psi[l][orb][r]
    in the files' order (missing l files are zero, as h_initialiseTDKSOSet.cuh
    loads l <= L0 only)."""
    psi = np.zeros((n_l, n_orb, N), dtype=np.complex128)
    for l in range(n_l):
        f = os.path.join(root, "eigenStates", "active", "%03d" % target_id, "eigenStatesL%03d.bin" % l)
        if os.path.exists(f):
            psi[l] = np.fromfile(f, dtype=np.complex128).reshape(n_orb, N)
    return psi


def params_macros(system, n_t, dt, n_scan, in_situ, extra=None):
    """The params.cuh macros the reference's analysis reads (analysis.py:
    111-143) and the hot path's sizes (init.py:493-604), as name -> value."""
    order, n_active = system.atom.reference_order()
    lact = [system.atom.orbitals[k][1] for k in order[:n_active]]
    mac = {
        "TD_SIMULATION": 1, "IN_SITU": int(bool(in_situ)),
        "N_r": system.N, "N_g": system.N + 1, "zeta": system.grid.zeta,
        "R_MAX": system.r_max, "N_l": system.n_l, "N_m": system.angular.n_m,
        "Z": system.atom.Z, "L0": int(max(lact)), "targetID": system.atom.target_id,
        "N_kso": n_active, "N_kso_ALL": system.n_orb,
        "N_t": int(n_t), "dt": float(dt), "N_scan": int(n_scan),
        "N_har": system.n_har, "USE_HARTREE": int(bool(system.interacting)),
    }
    if extra:
        mac.update(extra)
    return mac


def write_params(path, sim_date, index, macros):
    """params.cuh in the reference's form: '#define NAME (value)' lines
    (init.py:493-604), which its parseParams reads back as '(value)'."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    lines = ["// Simulation Parameters", "",
             "#define SIM_DATE (%d)" % int(sim_date), "#define SIM_INDEX (%d)" % int(index), ""]
    for k, v in macros.items():
        if isinstance(v, (bool, np.bool_)) or isinstance(v, (int, np.integer)):
            lines.append("#define %s (%d) " % (k, int(v)))
        else:
            lines.append("#define %s (%.15E) " % (k, float(v)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_output(root, index, system, time_grid, scans, sim_date=None, in_situ=False,
                 macros=None):
    """Write the run under root/DDDDDDDD/IIIII and return that directory.
    scans: list of dict(E=[N_t], chan=[N_t, n_orb, n_l] complex channel
    dipoles, acc_total=[N_t], omega, spectrum); an in-situ run holds N_scan
    perturbed scans and the unperturbed one last (main.cu:460-467), so its
    params.cuh N_scan is len(scans) - 1.  macros: further params.cuh values."""
    sim_date = sim_date_today() if sim_date is None else int(sim_date)
    base = os.path.join(root, "%08d" % sim_date, "%05d" % index)
    g = system.grid
    order, _ = system.atom.reference_order()
    nt = len(time_grid)
    n_scan = len(scans) - 1 if in_situ else len(scans)
    mac = params_macros(system, nt, time_grid[1] - time_grid[0] if nt > 1 else 0.0, n_scan,
                        in_situ, macros)
    write_params(os.path.join(base, "params.cuh"), sim_date, index, mac)
    _w(os.path.join(base, "static", "mesh", "r.bin"), g.r, np.float64)
    _w(os.path.join(base, "static", "mesh", "drdx.bin"), g.drdx, np.float64)
    _w(os.path.join(base, "static", "mesh", "time.bin"), time_grid, np.float64)
    td = os.path.join(base, "timeDependent")
    for s, sc in enumerate(scans):
        _w(os.path.join(td, "field", "E%03d.bin" % s), sc["E"], np.float64)
        chan = np.asarray(sc["chan"])
        if chan.shape != (nt, system.n_orb, system.n_l):
            raise ValueError("scan %d: channel dipoles of shape %s, expected %s"
                             % (s, chan.shape, (nt, system.n_orb, system.n_l)))
        for kk, k in enumerate(order):
            kdir = os.path.join(td, "observables", "dipoleMoment", "scan%03d" % s, "kso%03d" % kk)
            for l in range(system.n_l):
                _w(os.path.join(kdir, "dipL%03d.bin" % l), chan[:, k, l], np.complex128)
        _w(os.path.join(td, "observables", "dipoleAcceleration", "scan%03d" % s, "acc.bin"),
           sc["acc_total"], np.float64)
        if sc.get("spectrum") is not None:
            _w(os.path.join(td, "spectrum", "scan%03d" % s, "w.bin"), sc["omega"], np.float64)
            _w(os.path.join(td, "spectrum", "scan%03d" % s, "S.bin"), sc["spectrum"], np.float64)
    return base
