"""Write the FP64 oracle fixtures under tests/golden/:

he_c1_oracle.npz  the full helium run of the first configuration (N = 96
                  Gauss-Lobatto nodes to r_max = 60, Lmax = 12, Hartree to
                  L = 24, 800 nm sin^2 pulse of 2 cycles at 1e14 W/cm^2,
                  dt = 0.05, 4413 steps, imaginary-time ground state from
                  seed 0).  About 30 s on a CPU.
ar_c2_oracle.npz  2000 steps of argon at the second configuration's size
                  (N = 512 to r_max = 200, Lmax = 64, Hartree to L = 128,
                  dt = 0.02) from the ground state, driven by samples
                  20000..21999 of its 800 nm, 8-cycle, 2e14 W/cm^2 sin^2 pulse
                  (the pulse's peak, |E| up to E0: the field is switched on
                  suddenly at full strength).  About 7 minutes on a CPU.

Run from the repository root:
    python tests/golden/make_oracle_golden.py [he] [ar]"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import field as F, propagate as P  # noqa: E402

HE = ("He", 96, 60.0, 12, 0.05)
AR = ("Ar", 512, 200.0, 64, 0.02)
AR_WINDOW = (20000, 22000)
HERE = os.path.dirname(os.path.abspath(__file__))


def helium():
    S = P.System(*HE)
    w = F.omega_from_wavelength(800.0)
    E0 = F.e0_from_intensity(1e14)
    Em = F.midpoint_series(E0, w, 2, 0.05)
    res = S.run(Em)
    out = os.path.join(HERE, "he_c1_oracle.npz")
    np.savez_compressed(out, E_mid=Em, dip=res["dip"], acc=res["acc"], norm=res["norm"],
                        eps=S.gs["eps"], nkeep=np.array(S.nkeep))
    print("wrote", out, res["dip"].shape)


def argon():
    S = P.System(*AR)
    w = F.omega_from_wavelength(800.0)
    Em = F.midpoint_series(F.e0_from_intensity(2e14), w, 8, 0.02)[AR_WINDOW[0]:AR_WINDOW[1]]
    res = S.run(Em)
    out = os.path.join(HERE, "ar_c2_oracle.npz")
    np.savez_compressed(out, E_mid=Em, window=np.array(AR_WINDOW), dip=res["dip"],
                        acc=res["acc"], norm=res["norm"], eps=S.gs["eps"],
                        nkeep=np.array(S.nkeep), v_gs=S.v_gs)
    print("wrote", out, res["dip"].shape)


def main(argv):
    which = argv or ["he"]
    if "he" in which:
        helium()
    if "ar" in which:
        argon()


if __name__ == "__main__":
    main(sys.argv[1:])
