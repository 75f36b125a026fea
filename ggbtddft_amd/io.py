"""Reference-compatible files (raw little-endian, numpy ``tofile`` order).

Writers follow the layouts the reference's loaders and analysis read:

workingData/mesh/{r,P,drdx}.bin                f64 x N       (init.py:356-358 -> h_initialiseMesh.cuh)
workingData/mesh/time.bin                      f64 x N_t     (init.py:359)
workingData/eigenStates/active/TTT/eigenStatesLlll.bin
                                               c128 x (n_orb*N), [orb][r] per l (init.py:373-375)
workingData/eigenStates/active/TTT/eigenValues.bin  c128 x n_orb (init.py:371)
workingData/stateParameters/{l,m,occ}_active.bin    int32 x n_orb (init.py:377-379)
workingData/stateParameters/densityFrozen.bin  f64 x N       (init.py:380; zeros: all-electron)

output/IIIII/static/mesh/{r,drdx,time}.bin                 (h_writeMesh.cuh:1-18)
output/IIIII/timeDependent/field/ESSS.bin     f64 x N_t     (h_writeDipoleMoment.cuh:73-87)
output/IIIII/timeDependent/observables/dipoleMoment/scanSSS/ksoKKK/dipL000.bin
                                               c128 x N_t    (h_writeDipoleMoment.cuh:12-22)

The reference writes one dipole file per (orbital, l) channel; this
propagator reduces over l on the device, so it writes the per-orbital total
<r cos theta> into the l = 000 file and zero files for the other channels
(the reference's analysis sums the channels, analysis.py:411-432).  The
Ehrenfest acceleration and the spectrum are written beside them as
dipoleAcceleration/.../acc.bin (f64) and spectrum/scanSSS/{w,S}.bin (f64).
"""
import os

import numpy as np


def _w(path, arr, dtype):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.ascontiguousarray(arr, dtype=dtype).tofile(path)


def write_working_data(root, system, time_grid):
    g = system.grid
    _w(os.path.join(root, "mesh", "r.bin"), g.r, np.float64)
    _w(os.path.join(root, "mesh", "P.bin"), g.P, np.float64)
    _w(os.path.join(root, "mesh", "drdx.bin"), g.drdx, np.float64)
    _w(os.path.join(root, "mesh", "time.bin"), time_grid, np.float64)
    tid = "%03d" % system.atom.target_id
    for l in range(system.n_l):
        _w(os.path.join(root, "eigenStates", "active", tid, "eigenStatesL%03d.bin" % l),
           system.psi0[l], np.complex128)
    eps = np.array([system.gs.eps[o[4]] for o in system.atom.orbitals], dtype=np.complex128)
    _w(os.path.join(root, "eigenStates", "active", tid, "eigenValues.bin"), eps, np.complex128)
    sp = os.path.join(root, "stateParameters")
    _w(os.path.join(sp, "l_active.bin"), [o[1] for o in system.atom.orbitals], np.int32)
    _w(os.path.join(sp, "m_active.bin"), [o[2] for o in system.atom.orbitals], np.int32)
    _w(os.path.join(sp, "occ_active.bin"), [int(o[3]) for o in system.atom.orbitals], np.int32)
    _w(os.path.join(sp, "densityFrozen.bin"), np.zeros(system.N), np.float64)


def read_eigenstates(root, target_id, n_l, n_orb, N):
    """Load workingData/eigenStates/active/TTT/eigenStatesLlll.bin into psi[l][orb][r]
    (missing l files are zero, as h_initialiseTDKSOSet.cuh loads l <= L0 only)."""
    psi = np.zeros((n_l, n_orb, N), dtype=np.complex128)
    for l in range(n_l):
        f = os.path.join(root, "eigenStates", "active", "%03d" % target_id, "eigenStatesL%03d.bin" % l)
        if os.path.exists(f):
            psi[l] = np.fromfile(f, dtype=np.complex128).reshape(n_orb, N)
    return psi


def write_output(root, index, system, time_grid, scans):
    """scans: list of dict(E=[N_t], dip=[N_t, n_orb], acc_total=[N_t], omega, spectrum)."""
    base = os.path.join(root, "%05d" % index)
    g = system.grid
    _w(os.path.join(base, "static", "mesh", "r.bin"), g.r, np.float64)
    _w(os.path.join(base, "static", "mesh", "drdx.bin"), g.drdx, np.float64)
    _w(os.path.join(base, "static", "mesh", "time.bin"), time_grid, np.float64)
    td = os.path.join(base, "timeDependent")
    for s, sc in enumerate(scans):
        _w(os.path.join(td, "field", "E%03d.bin" % s), sc["E"], np.float64)
        dip = np.asarray(sc["dip"])
        for k in range(dip.shape[1]):
            kdir = os.path.join(td, "observables", "dipoleMoment", "scan%03d" % s, "kso%03d" % k)
            for l in range(system.n_l):
                val = dip[:, k] if l == 0 else np.zeros(dip.shape[0])
                _w(os.path.join(kdir, "dipL%03d.bin" % l), val.astype(np.complex128), np.complex128)
        _w(os.path.join(td, "observables", "dipoleAcceleration", "scan%03d" % s, "acc.bin"),
           sc["acc_total"], np.float64)
        if sc.get("spectrum") is not None:
            _w(os.path.join(td, "spectrum", "scan%03d" % s, "w.bin"), sc["omega"], np.float64)
            _w(os.path.join(td, "spectrum", "scan%03d" % s, "S.bin"), sc["spectrum"], np.float64)
    return base
