"""Atoms, orbital sets and unit conversions.

Electron configurations follow ``getElectronConfiguration``
(src/scripts/phys/internal.py:732-878): closed (n, l) shells split into
m = 0..l orbitals, occupation 2 for m = 0 and 4 for |m| > 0 (the +-m pair
at double weight).  Atom ids follow ``getTargetID`` (internal.py:675-694).
Unit constants follow init.py:10-14.
"""

ATTOSECOND_AU = 0.041341374575751
FEMTOSECOND_AU = 41.3413745758
INTENSITY_AU = 3.509e16
ELECTRONVOLT_AU = 27.2114
NM_TO_HARTREE = 45.5633525     # 1239.84193 eV nm / 27.2113862 eV

_SHELLS = {
    "H": (0, 1, ((1, 0),)),
    "He": (1, 2, ((1, 0),)),
    "Ne": (2, 10, ((1, 0), (2, 0), (2, 1))),
    "Ar": (3, 18, ((1, 0), (2, 0), (2, 1), (3, 0), (3, 1))),
    "Kr": (4, 36, ((1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1))),
    "Xe": (5, 54, ((1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1), (4, 2),
                   (5, 0), (5, 1))),
}
# The reference propagates only these valence orbitals (n, l, m) over a
# frozen core, in this order (getSimulationStates, internal.py:1162-1202),
# and its analysis script indexes its per-orbital files by that order
# (getKSOL/getKSOM, internal.py:16-42).
REFERENCE_ACTIVE = {
    "H": ((1, 0, 0),),
    "He": ((1, 0, 0),),
    "Ne": ((2, 0, 0), (2, 1, 0), (2, 1, 1)),
    "Ar": ((3, 0, 0), (3, 1, 0), (3, 1, 1)),
    "Kr": ((4, 0, 0), (4, 1, 0), (4, 1, 1)),
    "Xe": ((4, 0, 0), (5, 0, 0), (4, 1, 0), (4, 1, 1), (5, 1, 0), (5, 1, 1), (4, 2, 0), (4, 2, 1),
           (4, 2, 2)),
}
_NAMES = {"hydrogen": "H", "helium": "He", "neon": "Ne", "argon": "Ar", "krypton": "Kr",
          "xenon": "Xe"}
_L_LETTER = "spdfghik"


class Atom:
    def __init__(self, name):
        key = _NAMES.get(str(name).strip().lower())
        if key is None:
            for k in _SHELLS:
                if k.lower() == str(name).strip().lower():
                    key = k
        if key is None:
            raise ValueError("unsupported atom %r (H, He, Ne, Ar, Kr, Xe)" % (name,))
        self.symbol = key
        self.target_id, z, shells = _SHELLS[key]
        self.Z = float(z)
        # ground-state shells (n, l, occ)
        if key == "H":
            self.shells = [(1, 0, 1.0)]
        else:
            self.shells = [(n, l, 2.0 * (2 * l + 1)) for (n, l) in shells]
        # time-dependent orbitals grouped by m: (n, l, m, occ, shell index)
        orbs = []
        for s, (n, l, occ) in enumerate(self.shells):
            for m in range(l + 1):
                w = occ if key == "H" else (2.0 if m == 0 else 4.0)
                orbs.append((n, l, m, w, s))
        self.orbitals = sorted(orbs, key=lambda o: o[2])

    @property
    def labels(self):
        return ["%d%s%d" % (n, _L_LETTER[l], m) for (n, l, m, _, _) in self.orbitals]

    @property
    def n_electrons(self):
        return sum(o for (_, _, o) in self.shells)

    def reference_order(self):
        """(order, n_active): the orbital indices with the reference's active
        orbitals first, in its order, then the remaining (core) orbitals in
        this propagator's order; n_active = how many are the reference's."""
        key = [(n, l, m) for (n, l, m, _, _) in self.orbitals]
        active = [key.index(o) for o in REFERENCE_ACTIVE[self.symbol]]
        rest = [k for k in range(len(key)) if k not in active]
        return active + rest, len(active)
