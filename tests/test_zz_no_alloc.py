"""pyseqm_fock_noalloc (library-owned scratch, no allocation, no scratch argument) against
pyseqm_fock (caller scratch) through the C ABI: the same kernels in the same order, so F and
Eelec are bitwise identical.  Two calls in a row on the tile-parallel kernels that use the
scratch, the second on a smaller batch (the scratch keeps the first call's contents); then a
batch of small molecules (the per-molecule kernel, no scratch), plus an active-molecule mask."""
import numpy as np
import pytest
import torch

from pyseqm_amd import _native as nat
from pyseqm_amd import workloads as W
from pyseqm_amd.engine import SCFEngine

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _random_density(eng, seed):
    b = eng.b
    g = torch.Generator(device=DEV).manual_seed(seed)
    P = eng.initial_density()
    for m in range(min(b.nmol, 4)):
        n, o = int(b.norb[m]), int(b.mol_mat_off[m])
        x = torch.randn(n, n, generator=g, device=DEV, dtype=torch.float64) * 0.1
        P[o:o + n * n] += (x + x.T).reshape(-1)
    return P


def _compare(eng, P, active=None):
    b = eng.b
    F1 = torch.empty_like(P)
    e1 = torch.zeros(b.nmol, dtype=torch.float64, device=DEV)
    eng.fock(P, F1, active=active, eelec=e1)
    F2 = torch.full_like(P, float("nan")) if active is None else F1.clone()
    e2 = torch.zeros(b.nmol, dtype=torch.float64, device=DEV)
    need = nat.lib().pyseqm_fock_scratch_doubles(nat.ctypes.byref(b.L))
    assert 0 < need <= nat.FOCK_SCRATCH_DOUBLES
    print(f"batch {b.nmol} molecules, {b.natoms} atoms: scratch {need * 8 / 2**20:.1f} MiB", flush=True)
    nat.call("pyseqm_fock_noalloc", b.L, nat.ptr(eng.atom_par), nat.ptr(eng.w_xx), nat.ptr(eng.w_xh),
             nat.ptr(eng.w_hh), nat.ptr(P), nat.ptr(eng.H), nat.ptr(F2), nat.ptr(e2), nat.ptr(active), eng.stream)
    torch.cuda.synchronize()
    assert torch.equal(F1, F2)
    assert torch.equal(e1, e2)


def test_fock_noalloc_equals_fock_on_two_calls_the_second_smaller():
    sp, x = W.water_clusters(1024, 1, seed=0)  # 3072 atoms, 6144 orbitals: 37 of the 48 MiB
    big = SCFEngine(sp, x, "PM3", DEV)
    _compare(big, _random_density(big, 1))
    sp, x = W.water_clusters(64, 2, seed=1)  # 2 x 192 atoms: still above the per-molecule kernel's 128
    mid = SCFEngine(sp, x, "PM3", DEV)
    assert mid.b.L.max_mol_atoms > 128 and mid.b.natoms < big.b.natoms
    _compare(mid, _random_density(mid, 3))
    sp, x = W.alkanes(16, seed=3)
    small = SCFEngine(sp, x, "AM1", DEV)
    P = _random_density(small, 2)
    _compare(small, P)
    act = torch.as_tensor((np.arange(16) % 3 != 0).astype(np.uint8), device=DEV)
    _compare(small, P, active=act)
