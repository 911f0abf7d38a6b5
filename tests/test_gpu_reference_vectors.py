"""The GPU engine against energies written by the reference itself (tests/golden/reference_vectors.json,
made by tests/golden/make_reference_vectors.py from lanl/PYSEQM on the CPU).

Bars: FP64 density (diagonalisation) Hf / Etot / Eelec / Enuc within 1e-6 eV of the reference and the
same SCF iteration count within one; the mixed-precision SP2 density within 1e-4 eV/atom.  The 128
alkanes are the first 128 of the 1024-alkane benchmark batch: the test also asserts that this run
takes the benchmark's density path (one launch of one-workgroup-per-molecule eigensolves over
every block, largest first) and that it holds atom pairs beyond the 40-bohr overlap cutoff
(hcore.py:83).
"""
import numpy as np
import pytest
import torch

from pyseqm_amd import engine as E
from pyseqm_amd.engine import SCFEngine
from tests.test_reference_vectors import CASE_NAMES, CASES, case_input

pytestmark = pytest.mark.gpu
DEV = "cuda"
A0 = 0.529167


def _run(c, density="eigh", eps=None):
    sp, x = case_input(c)
    eng = SCFEngine(sp, x, c["method"], DEV)
    r = eng.scf(eps=c["eps"] if eps is None else eps, converger=tuple(c["converger"]), density=density)
    en = eng.energies(r.P)
    torch.cuda.synchronize()
    return eng, r, {k: v.cpu().numpy() for k, v in en.items() if k != "F"}, sp, x


@pytest.mark.parametrize("name", CASE_NAMES)
def test_fp64_engine_matches_reference(name):
    c = CASES[name]
    print(f"case {name} ({c['method']}, converger {c['converger']})", flush=True)
    eng, r, en, sp, x = _run(c)
    assert not r.notconverged.any()
    for k in ("Hf", "Etot", "Eelec", "Enuc"):
        np.testing.assert_allclose(en[k], c[k], rtol=0, atol=1e-6, err_msg=f"{name} {k}")
    assert abs(r.iterations - c["scf_iterations"]) <= 1, (r.iterations, c["scf_iterations"])


def test_c2_slice_takes_the_benchmark_path():
    """First 128 of the 1024 alkanes: one workgroup-eigensolver launch per SCF iteration, and
    >40-bohr pairs."""
    c = CASES["am1_alkanes_c2_first128"]
    sp, x = case_input(c)
    # atom pairs beyond the overlap cutoff (their overlap, hence beta term, is zero)
    far = 0
    for m in range(sp.shape[0]):
        n = int((sp[m] > 0).sum())
        d = np.linalg.norm(x[m, :n, None] - x[m, None, :n], axis=-1) / A0
        far += int((d > 40.0).sum())
    assert far > 0
    eng, r, en, _, _ = _run(c)
    first = eng.eigh_log[0]  # the first SCF iteration solves every block in one launch
    assert first == (eng.b.nmol, int(eng.b.norb.max()), "workgroup"), first
    assert all(kind == "workgroup" for _, _, kind in eng.eigh_log)
    assert int(eng.eigh_fail.item()) == 0
    np.testing.assert_allclose(en["Hf"], c["Hf"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("name", ["am1_alkanes_c2_first128", "pm3_water64", "pm3_benzene16_c1"])
def test_mixed_sp2_within_1e4_per_atom_of_reference(name):
    c = CASES[name]
    eng, r, en, sp, _ = _run(c, density="sp2", eps=1e-5)
    assert not r.notconverged.any()
    nat = (sp > 0).sum(1)
    err = np.abs(en["Hf"] - np.array(c["Hf"])) / nat
    print(f"{name}: SP2 |dHf|/atom max {err.max():.2e} mean {err.mean():.2e} eV", flush=True)
    assert err.max() <= 1e-4
