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


def test_orbital_energy_gradient_matches_reconverged_slope():
    """scf_loop(eig=True) with scf_backward = 1: the orbital energies carry the implicit gradient
    too (e_k = c_k^T F(theta, P*) c_k).  The loss, sum of the occupied energies plus 0.05 sum e_k^2
    plus a density term, is a symmetric function of each degenerate level's energies, so it stays
    smooth where benzene's and NH3's levels are degenerate."""
    sp, c = _batch()
    p = {"method": "AM1", "scf_eps": 1e-10, "scf_converger": [2]}
    const = pyseqm_amd.Constants().to(DEV)
    mol = pyseqm_amd.Molecule(const, p, torch.as_tensor(c, device=DEV), torch.as_tensor(sp, device=DEV))
    eng = mol.engine()
    name = "U_pp"
    col = COLUMNS.index(name)
    base = eng.atom_par[:, col].clone()
    n4 = 4 * mol.molsize
    g = torch.Generator(device=DEV).manual_seed(5)
    W = torch.randn(mol.nmol, n4, n4, generator=g, device=DEV, dtype=torch.float64)
    W = (W + W.transpose(1, 2)) / (4.0 * n4)
    occ = torch.arange(n4, device=DEV)[None, :] < mol.nocc.reshape(-1, 1)

    def loss(out):
        e, P = out[1], out[2]
        return (e * occ).sum() + 0.05 * (e * e).sum() + (W * P).sum()

    theta = base.clone().requires_grad_(True)
    mol.use_parameters({name: theta})
    out = scf_loop(mol, eps=1e-10, scf_converger=[2], eig=True, scf_backward=1, scf_backward_eps=1e-11)
    assert out[1].requires_grad and out[2].requires_grad
    loss(out).backward()
    grad = theta.grad.detach().clone()
    scale = float(grad.abs().max())
    Z = np.asarray(eng.b.Z)
    for a in [int(np.nonzero(Z == z)[0][0]) for z in sorted(set(Z.tolist())) if z > 1]:
        h = 1e-3 * max(abs(float(base[a])), 0.1)
        vals = []
        for s in (-2, -1, 1, 2):
            v = base.clone()
            v[a] += s * h
            mol.use_parameters({name: v})
            with torch.no_grad():
                vals.append(float(loss(scf_loop(mol, eps=1e-10, scf_converger=[2], eig=True))))
        d = (vals[0] - 8.0 * vals[1] + 8.0 * vals[2] - vals[3]) / (12.0 * h)
        print(f"U_pp atom {a} (Z={Z[a]}): autograd {float(grad[a]):.9e} slope {d:.9e}", flush=True)
        assert abs(float(grad[a]) - d) <= 2e-5 * scale, (a, float(grad[a]), d, scale)


def test_implicit_backward_above_the_workgroup_limit():
    """(H2O)_32 (192 orbitals, past the workgroup eigensolver's 139-orbital all-eigenpairs limit: its
    eigenpairs come from the batched library eigensolver) beside methanal (workgroup solver): the
    implicit gradient of the weighted-density loss against the re-converged slope, PM3 g_ss."""
    from pyseqm_amd.workloads import _pad, water_cluster
    from tests.conftest import read_xyz
    sp, c = _pad([water_cluster(32, seed=1), tuple(a[0] for a in read_xyz(["methanal.1.xyz"]))])
    p = {"method": "PM3", "scf_eps": 1e-10, "scf_converger": [2]}
    const = pyseqm_amd.Constants().to(DEV)
    mol = pyseqm_amd.Molecule(const, p, torch.as_tensor(c, device=DEV), torch.as_tensor(sp, device=DEV))
    eng = mol.engine()
    assert int(eng.b.norb.max()) == 192
    name = "g_ss"
    base = eng.atom_par[:, COLUMNS.index(name)].clone()
    g = torch.Generator(device=DEV).manual_seed(3)
    n4 = 4 * mol.molsize
    W = torch.randn(mol.nmol, n4, n4, generator=g, device=DEV, dtype=torch.float64)
    W = (W + W.transpose(1, 2)) / (2.0 * n4)
    theta = base.clone().requires_grad_(True)
    mol.use_parameters({name: theta})
    print("(H2O)_32 + methanal PM3: implicit backward", flush=True)
    (W * _density(mol, 1e-10, 1)).sum().backward()
    grad = theta.grad.detach().clone()
    scale = float(grad.abs().max())
    assert torch.isfinite(grad).all() and scale > 0.0
    Z = np.asarray(eng.b.Z)
    # the first O and H of the cluster, methanal's C
    for a in [0, 32, int(np.nonzero(Z == 6)[0][0])]:
        h = 1e-3 * max(abs(float(base[a])), 0.1)
        vals = []
        for s in (-2, -1, 1, 2):
            v = base.clone()
            v[a] += s * h
            mol.use_parameters({name: v})
            with torch.no_grad():
                vals.append(float((W * _density(mol, 1e-10, 0)).sum()))
        d = (vals[0] - 8.0 * vals[1] + 8.0 * vals[2] - vals[3]) / (12.0 * h)
        print(f"atom {a} (Z={Z[a]}): autograd {float(grad[a]):.9e} slope {d:.9e}", flush=True)
        assert abs(float(grad[a]) - d) <= 2e-5 * scale, (a, float(grad[a]), d, scale)
