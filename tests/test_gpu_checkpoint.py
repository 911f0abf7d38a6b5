"""Checkpoint / restore of device-resident XL-BOMD state (MolecularDynamics.py:1099-1166): a run
resumed from a checkpoint continues exactly as the uninterrupted run (coordinates bitwise, energies
to 1e-12 eV), both for the drop-in XL_BOMD driver and for the serving loop's graph-captured batch."""
import numpy as np
import pytest
import torch

import pyseqm_amd
from pyseqm_amd import workloads as W
from pyseqm_amd.serving import XLService
from tests.conftest import read_xyz

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("cls,params", [("XL_BOMD", {"k": 6}),
                                        ("KSA_XL_BOMD", {"k": 6, "max_rank": 3, "err_threshold": 0.0, "T_el": 1500})])
def test_xl_bomd_run_from_checkpoint_continues_the_run(tmp_path, cls, params):
    """Also for KSA_XL_BOMD (the reference's test_md_checkpoint_resume.py:200-230 covers both): the
    Krylov kernel dP2dt2 and the entropy term travel in the checkpoint."""
    sp, c = read_xyz(["methanal.1.xyz", "methane.xyz"])
    p = {"method": "AM1", "scf_eps": 1.0e-7, "scf_converger": [1]}

    def md_and_mol():
        torch.manual_seed(0)
        const = pyseqm_amd.Constants().to(DEV)
        mol = pyseqm_amd.Molecule(const, dict(p), torch.as_tensor(c, device=DEV), torch.as_tensor(sp, device=DEV))
        md = getattr(pyseqm_amd, cls)(xl_bomd_params=dict(params), seqm_parameters=dict(p), timestep=0.5, Temp=300.0)
        return md, mol

    md, mol = md_and_mol()
    md.run(mol, steps=8, seed=0)
    T, Ek, Ep, Etot = md.thermo()
    x_full = mol.coordinates.detach().clone()

    md1, mol1 = md_and_mol()
    md1.run(mol1, steps=5, seed=0)
    path = str(tmp_path / "xl.safetensors")
    md1.save_checkpoint(mol1, 8, step_done=5, path=path)
    md2, mol2 = getattr(pyseqm_amd, cls).run_from_checkpoint(path, device=DEV)
    assert type(md2).__name__ == cls
    T2, Ek2, Ep2, Etot2 = md2.thermo()
    assert torch.equal(mol2.coordinates.detach(), x_full)
    np.testing.assert_allclose(Etot2, Etot, atol=1e-12, rtol=0)


def test_xl_service_restore_replays_the_same_steps(tmp_path):
    sp, c = W.alkanes(6, seed=2, nmax=8)
    svc = XLService(sp, c, "AM1", DEV, k=6, graphs=True)
    for _ in range(3):
        svc.step()
    path = str(tmp_path / "svc.safetensors")
    svc.checkpoint(path)
    xs = []
    for _ in range(9):  # across a full ring of captured graphs
        svc.step()
        xs.append(svc.traj.x.clone())
    e_first = svc.energies()
    svc.restore(path)
    for k in range(9):
        svc.step()
        assert torch.equal(svc.traj.x, xs[k]), k
    np.testing.assert_array_equal(svc.energies(), e_first)
