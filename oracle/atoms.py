"""All-electron orbital sets (oracle; test infrastructure only).

Shell lists restate ``internal.py:732-878`` (getElectronConfiguration):
each closed (n, l) shell becomes the m = 0..l orbitals with occupation 2 for
m = 0 and 4 for m > 0 (the +-m pair carries double weight, ``internal.py:764-772``).
The time-dependent orbitals are stable-sorted by m so every |m| group is
contiguous.  Ground-state shells (one radial function per (n, l)) carry
occupation 2(2l+1); hydrogen carries 1.
"""

ATOMS = {
    # symbol: (Z, [(n, l), ...] in internal.py order)
    "H": (1, [(1, 0)]),
    "He": (2, [(1, 0)]),
    "Ne": (10, [(1, 0), (2, 0), (2, 1)]),
    "Ar": (18, [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1)]),
    "Kr": (36, [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1)]),
    "Xe": (54, [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1),
                (4, 2), (5, 0), (5, 1)]),
}

ALIASES = {"hydrogen": "H", "helium": "He", "neon": "Ne", "argon": "Ar",
           "krypton": "Kr", "xenon": "Xe"}


def canonical(symbol):
    s = symbol.strip()
    if s.lower() in ALIASES:
        return ALIASES[s.lower()]
    for k in ATOMS:
        if k.lower() == s.lower():
            return k
    raise ValueError("unknown atom %r" % symbol)


def shells(symbol):
    """[(n, l, occ)] ground-state shells."""
    sym = canonical(symbol)
    Z, sh = ATOMS[sym]
    if sym == "H":
        return [(1, 0, 1.0)]
    return [(n, l, 2.0 * (2 * l + 1)) for (n, l) in sh]


def td_orbitals(symbol):
    """[(n, l, m, occ, shell_index)] sorted by m (stable)."""
    sym = canonical(symbol)
    Z, sh = ATOMS[sym]
    orbs = []
    for s, (n, l) in enumerate(sh):
        for m in range(0, l + 1):
            occ = 1.0 if sym == "H" else (2.0 if m == 0 else 4.0)
            orbs.append((n, l, m, occ, s))
    orbs.sort(key=lambda o: o[2])
    return orbs


def nuclear_charge(symbol):
    return float(ATOMS[canonical(symbol)][0])
