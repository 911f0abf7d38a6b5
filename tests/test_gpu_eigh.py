"""The one-workgroup-per-molecule FP64 eigensolver (pyseqm_eigh_density) on the GPU.

Bars: P = 2 C_occ C_occ^T within 1e-10 of numpy's float64 eigh density of the same Fock blocks
(measured ~1e-14), exactly symmetric, trace 2 nocc and idempotent to 1e-10; the same result on a
rerun (bitwise); no vector lost to Gram-Schmidt, also for blocks with exactly degenerate occupied levels.
"""
import numpy as np
import pytest
import torch

from pyseqm_amd import workloads as W
from pyseqm_amd.engine import SCFEngine, eigh_wg_ok
from tests.conftest import read_xyz

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _blocks(eng, X):
    b = eng.b
    out = []
    for m in range(b.nmol):
        n, o = int(b.norb[m]), int(b.mol_mat_off[m])
        out.append(X[o:o + n * n].view(n, n).cpu().numpy())
    return out


def _reference_density(F, nocc):
    w, U = np.linalg.eigh(F)
    return 2.0 * U[:, :nocc] @ U[:, :nocc].T


def _run_wg(eng, F):
    b = eng.b
    ids = np.arange(b.nmol)
    assert eigh_wg_ok(b.norb, b.nocc).all()
    P = torch.zeros_like(F)
    eng._eigh_wg(F, ids, P)
    torch.cuda.synchronize()
    return P


def _check(eng, F, P, tol=1e-10):
    b = eng.b
    for m, (f, p) in enumerate(zip(_blocks(eng, F), _blocks(eng, P))):
        nocc = int(b.nocc[m])
        ref = _reference_density(f, nocc)
        err = np.abs(p - ref).max()
        assert err < tol, (m, f.shape[0], err)
        assert np.array_equal(p, p.T), m
        assert abs(np.trace(p) - 2 * nocc) < tol, m
        assert np.abs(p @ p - 2 * p).max() < tol, m
    assert int(eng.eigh_fail.item()) == 0


@pytest.mark.parametrize("name", ["alkanes", "methane+benzene", "water8", "benzene"])
def test_workgroup_eigh_matches_numpy(name):
    if name == "alkanes":
        sp, x = W.alkanes(64, seed=2)  # 14-182 orbitals
        method = "AM1"
    elif name == "methane+benzene":
        sp, x = read_xyz(["methane.xyz", "benzene.xyz"])
        method = "AM1"
    elif name == "water8":
        sp, x = W.water_clusters(8, 3, seed=1)
        method = "PM3"
    else:
        sp, x = W.benzene(16, seed=0)
        method = "PM3"
    eng = SCFEngine(sp, x, method, DEV)
    P0 = eng.initial_density()
    F = torch.empty_like(P0)
    eng.fock(P0, F)
    P = _run_wg(eng, F)
    _check(eng, F, P)
    P2 = _run_wg(eng, F)
    assert torch.equal(P, P2)


def test_workgroup_eigh_exact_degeneracies():
    """Fock blocks Q diag(lambda) Q^T whose occupied (and virtual) levels come in exactly degenerate
    triples: the occupied subspace is still unique, and the solver's vectors must span it."""
    sp, x = W.alkanes(6, seed=4, nmin=3, nmax=25)
    eng = SCFEngine(sp, x, "AM1", DEV)
    b = eng.b
    F = torch.zeros(b.total, dtype=torch.float64, device=DEV)
    rng = np.random.default_rng(0)
    for m in range(b.nmol):
        n, o, nocc = int(b.norb[m]), int(b.mol_mat_off[m]), int(b.nocc[m])
        occ = np.repeat(np.sort(rng.uniform(-40.0, -10.0, size=(nocc + 2) // 3)), 3)[:nocc]
        vir = np.repeat(np.sort(rng.uniform(1.0, 20.0, size=(n - nocc + 2) // 3)), 3)[:n - nocc]
        lam = np.concatenate([occ, vir])
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        f = Q @ np.diag(lam) @ Q.T
        f = 0.5 * (f + f.T)
        F[o:o + n * n] = torch.as_tensor(f.reshape(-1), device=DEV)
    P = _run_wg(eng, F)
    _check(eng, F, P, tol=1e-9)


def test_scf_density_path_uses_the_workgroup_solver():
    sp, x = W.alkanes(48, seed=6)
    eng = SCFEngine(sp, x, "AM1", DEV)
    r = eng.scf(eps=1e-8, converger=(2,))
    assert not r.notconverged.any()
    kinds = {k for _, _, k in eng.eigh_log}
    assert kinds == {"workgroup"}, kinds
    assert int(eng.eigh_fail.item()) == 0
