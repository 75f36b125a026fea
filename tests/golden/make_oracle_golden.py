"""Write tests/golden/he_c1_oracle.npz: the FP64 oracle's full helium run of
the first configuration (N = 96 Gauss-Lobatto nodes to r_max = 60, Lmax = 12,
Hartree to L = 24, 800 nm sin^2 pulse of 2 cycles at 1e14 W/cm^2, dt = 0.05,
4413 steps, imaginary-time ground state from seed 0).  About 30 s on a CPU.
Run from the repository root: python tests/golden/make_oracle_golden.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import field as F, propagate as P  # noqa: E402

ARGS = ("He", 96, 60.0, 12, 0.05)


def main():
    S = P.System(*ARGS)
    w = F.omega_from_wavelength(800.0)
    E0 = F.e0_from_intensity(1e14)
    Em = F.midpoint_series(E0, w, 2, 0.05)
    res = S.run(Em)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "he_c1_oracle.npz")
    np.savez_compressed(out, E_mid=Em, dip=res["dip"], acc=res["acc"], norm=res["norm"],
                        eps=S.gs["eps"], nkeep=np.array(S.nkeep))
    print("wrote", out, res["dip"].shape)


if __name__ == "__main__":
    main()
