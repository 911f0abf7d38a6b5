"""SCF implicit backward (scf_backward = 1, scf_loop.py:1779-1993) on the device: the gradient of a
density-dependent loss L(P*) with respect to learned per-atom parameters, through the converged
density, against the slope of the same loss over SCFs re-converged at theta +- h (4th-order
central differences, FP64 eigensolver, SCF eps 1e-10).  The loss weights the padded density with
a fixed random symmetric matrix, so every element of the adjoint reaches it.  Bar: 2e-5 relative
to the largest gradient entry of the batch (the re-converged densities carry ~1e-9 of SCF noise).
The adjoint solve, the density response (eigenpairs, eigenbasis products, the 2 / (e_i - e_a)
weights), the two-electron operator and the parameter contraction are all pinned by this slope."""
import numpy as np
import pytest
import torch

import pyseqm_amd
from pyseqm_amd.seqm_functions.parameters import COLUMNS
from pyseqm_amd.seqm_functions.scf_loop import scf_loop
from tests.test_gpu_param_grad import _batch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _density(mol, eps, backward):
    return scf_loop(mol, eps=eps, scf_converger=[2], scf_backward=backward, scf_backward_eps=1e-11)[2]


@pytest.mark.parametrize("method,name", [("AM1", "g_ss"), ("AM1", "zeta_s"), ("PM3", "beta_p"), ("MNDO", "h_sp")])
def test_implicit_backward_matches_reconverged_slope(method, name):
    sp, c = _batch()
    p = {"method": method, "scf_eps": 1e-10, "scf_converger": [2]}
    const = pyseqm_amd.Constants().to(DEV)
    mol = pyseqm_amd.Molecule(const, p, torch.as_tensor(c, device=DEV), torch.as_tensor(sp, device=DEV))
    eng = mol.engine()
    col = COLUMNS.index(name)
    base = eng.atom_par[:, col].clone()
    g = torch.Generator(device=DEV).manual_seed(11)
    n4 = 4 * mol.molsize
    W = torch.randn(mol.nmol, n4, n4, generator=g, device=DEV, dtype=torch.float64)
    W = (W + W.transpose(1, 2)) / (2.0 * n4)
    theta = base.clone().requires_grad_(True)
    mol.use_parameters({name: theta})
    P = _density(mol, 1e-10, 1)
    assert P.requires_grad
    (W * P).sum().backward()
    grad = theta.grad.detach().clone()
    assert torch.isfinite(grad).all()
    scale = float(grad.abs().max())
    assert scale > 0.0
    Z = np.asarray(eng.b.Z)
    # one atom of each element present
    atoms = [int(np.nonzero(Z == z)[0][0]) for z in sorted(set(Z.tolist()))]
    for a in atoms:
        h = 1e-3 * max(abs(float(base[a])), 0.1)
        vals = []
        for s in (-2, -1, 1, 2):
            v = base.clone()
            v[a] += s * h
            mol.use_parameters({name: v})
            with torch.no_grad():
                vals.append(float((W * _density(mol, 1e-10, 0)).sum()))
        d = (vals[0] - 8.0 * vals[1] + 8.0 * vals[2] - vals[3]) / (12.0 * h)
        print(f"{method} {name} atom {a} (Z={Z[a]}): autograd {float(grad[a]):.9e} slope {d:.9e}", flush=True)
        assert abs(float(grad[a]) - d) <= 2e-5 * scale, (a, float(grad[a]), d, scale)


def test_scf_backward_zero_leaves_the_density_without_graph():
    sp, c = _batch()
    p = {"method": "AM1", "scf_eps": 1e-8, "scf_converger": [2]}
    const = pyseqm_amd.Constants().to(DEV)
    mol = pyseqm_amd.Molecule(const, p, torch.as_tensor(c, device=DEV), torch.as_tensor(sp, device=DEV))
    theta = mol.engine().atom_par[:, COLUMNS.index("g_ss")].clone().requires_grad_(True)
    mol.use_parameters({"g_ss": theta})
    assert not _density(mol, 1e-8, 0).requires_grad
    with pytest.raises(ValueError):
        scf_loop(mol, eps=1e-8, scf_converger=[2], scf_backward=3)
