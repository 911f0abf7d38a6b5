import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
_EL = {"H": 1, "C": 6, "N": 7, "O": 8, "S": 16}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


def read_xyz(names):
    """Padded species/coordinates from tests/golden/data/*.xyz, atoms sorted by decreasing Z
    (the reference's read_xyz ordering, seqm/seqm_functions/read_xyz.py:16-52)."""
    mols = []
    for nm in names:
        lines = open(os.path.join(GOLDEN, "data", nm)).read().split("\n")
        n = int(lines[0])
        d = []
        for ln in lines[2:2 + n]:
            a, *x = ln.split()
            d.append([_EL[a]] + [float(v) for v in x])
        d = np.array(d)
        mols.append(d[np.argsort(-d[:, 0], kind="stable")])
    K = max(len(m) for m in mols)
    sp = np.zeros((len(mols), K), dtype=np.int64)
    xyz = np.zeros((len(mols), K, 3))
    for i, m in enumerate(mols):
        sp[i, :len(m)] = m[:, 0]
        xyz[i, :len(m)] = m[:, 1:]
    return sp, xyz


def golden(name):
    import json
    with open(os.path.join(GOLDEN, "reference", name + ".json")) as fh:
        return json.load(fh)


@pytest.fixture
def xyz():
    return read_xyz
