"""read_xyz: .xyz files -> padded species / coordinates (seqm/seqm_functions/read_xyz.py:16-52).

Each file holds an atom count, a comment line and one line per atom (element symbol or atomic
number, then x y z in Angstrom).  Atoms are sorted by decreasing atomic number (stable), the
order the batch layout requires, and the molecules are zero-padded to the largest atom count.
"""
import numpy as np

_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br "
            "Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe").split()
ELEMENTS = {s: z for z, s in enumerate(_SYMBOLS, start=1)}


def _atoms(fn):
    with open(fn) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0].split()[0])
    Z = np.zeros(n, dtype=np.int64)
    X = np.zeros((n, 3))
    for k, line in enumerate(lines[2:2 + n]):
        t = line.split()
        Z[k] = int(t[0]) if t[0].isdigit() else ELEMENTS[t[0]]
        X[k] = [float(v) for v in t[1:4]]
    return Z, X


def read_xyz(files, sort=True):
    """Returns species (M, K) int64 and coordinates (M, K, 3) float64 for M files."""
    mols = []
    for fn in files:
        Z, X = _atoms(fn)
        if sort:
            o = np.argsort(-Z, kind="stable")
            Z, X = Z[o], X[o]
        mols.append((Z, X))
    K = max((len(Z) for Z, _ in mols), default=0)
    species = np.zeros((len(mols), K), dtype=np.int64)
    coords = np.zeros((len(mols), K, 3))
    for i, (Z, X) in enumerate(mols):
        species[i, :len(Z)] = Z
        coords[i, :len(Z)] = X
    return species, coords
