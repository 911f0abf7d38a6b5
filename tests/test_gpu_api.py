"""The drop-in API (Molecule / Electronic_Structure / scf_loop) on the GPU against the
reference goldens, in the reference's padded layouts."""
import numpy as np
import pytest
import torch

import pyseqm_amd
from pyseqm_amd.seqm_functions.scf_loop import scf_loop
from tests.conftest import golden, read_xyz

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _molecule(files, p):
    sp, c = read_xyz(files)
    const = pyseqm_amd.Constants().to(DEV)
    return pyseqm_amd.Molecule(const, p, torch.as_tensor(c, device=DEV), torch.as_tensor(sp, device=DEV))


@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
def test_electronic_structure_methane_golden(method):
    p = {"method": method, "scf_eps": 1.0e-6, "scf_converger": [1]}
    mol = _molecule(["methane.xyz"], p)
    pyseqm_amd.Electronic_Structure(p)(mol)
    g = golden(f"smoke_single_point_{method}")
    for k in ("Etot", "Eelec", "Enuc"):
        assert abs(getattr(mol, k).item() - g[k]) < 1e-9, (k, getattr(mol, k).item(), g[k])
    np.testing.assert_allclose(mol.force.cpu().numpy().reshape(-1), np.array(g["force"]).reshape(-1), atol=1e-6)
    assert mol.dm.shape == (1, 4 * mol.molsize, 4 * mol.molsize)
    tr = mol.dm.diagonal(dim1=1, dim2=2).sum(1)
    assert abs(tr.item() - 2 * mol.nocc.item()) < 1e-9
    assert torch.all(mol.notconverged == 0) if torch.is_tensor(mol.notconverged) else True


def test_electronic_structure_batch_golden():
    p = {"method": "AM1", "scf_eps": 1.0e-7, "scf_converger": [1]}
    mol = _molecule(["methane.xyz", "benzene.xyz"], p)
    pyseqm_amd.Electronic_Structure(p)(mol)
    np.testing.assert_allclose(mol.Etot.cpu().numpy(), golden("batch_single_point_am1")["Etot"], atol=1e-8)
    f = mol.force.cpu().numpy()
    np.testing.assert_allclose(f.reshape(-1), np.array(golden("ground_force_batch_mixed")["force"]).reshape(-1),
                               atol=1e-6)
    # charges: tore - per-atom diagonal sums, neutral molecules
    assert np.abs(mol.q.sum(dim=1).cpu().numpy()).max() < 1e-8
    e_gap = mol.e_gap.cpu().numpy()
    assert np.all(e_gap > 0)


def test_scf_loop_returns_reference_tuple():
    p = {"method": "PM3", "scf_eps": 1.0e-8, "scf_converger": [2]}
    mol = _molecule(["methanal.1.xyz", "benzene.xyz"], p)
    F, e, P, Hcore, w, charge, rho0xi, rho0xj, riXH, ri, nc, v = scf_loop(mol, eps=1e-8, scf_converger=[2], eig=True)
    n4 = 4 * mol.molsize
    assert F.shape == P.shape == Hcore.shape == (2, n4, n4)
    assert w.shape == (mol.idxi.shape[0], 10, 10)
    assert e.shape[0] == 2 and charge.shape == (2, n4, mol.molsize)
    # symmetric F and P; P idempotent per molecule; charge rows of occupied orbitals sum to 1
    assert torch.allclose(F, F.transpose(1, 2), atol=1e-12)
    for m in range(2):
        n = int(mol.norb[m])
        Pm = P[m]
        assert torch.allclose(Pm @ Pm, 2 * Pm, atol=1e-7)
        np.testing.assert_allclose(charge[m, :n].sum(1).cpu().numpy(), 1.0, atol=1e-10)
    assert not nc.any()
