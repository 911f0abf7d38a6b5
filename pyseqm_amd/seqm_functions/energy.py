"""Energy terms with the reference's signatures (seqm/seqm_functions/energy.py): electronic
energies on the per-molecule reduction kernel, core-core repulsion on the pair kernel, and the
per-molecule sums / heat of formation."""
import numpy as np
import torch

from .. import _native as nat
from ..layout import MatrixLayout
from .constants import TORE

_METHODS = {"MNDO": 0, "AM1": 1, "PM3": 2}


def _sym(Hcore, doTriu):
    return Hcore.triu() + Hcore.triu(1).transpose(1, 2) if doTriu else Hcore


def _xl_energy(D, P, F, H):
    dev = nat.require_device(P.device)
    nmol, n, _ = P.shape
    b = MatrixLayout(nmol, n, 0, dev)
    out = torch.empty(nmol, dtype=torch.float64, device=dev)
    red = torch.empty(nmol * nat.MAX_RED_BLOCKS, dtype=torch.float64, device=dev)
    c = lambda t: t.to(torch.float64).contiguous()
    D, P, F, H = c(D), c(P), c(F), c(H)
    nat.call("pyseqm_xl_energy", b.L, nat.ptr(D), nat.ptr(P), nat.ptr(F), nat.ptr(H), nat.ptr(out), nat.ptr(red),
             nat.red_blocks(n), nat.stream_ptr(dev))
    return out


def elec_energy(P, F, Hcore, doTriu=True):
    """0.5 sum P (H + F) per molecule (energy.py:26-49), closed shell."""
    if F.dim() == 4:
        raise NotImplementedError("the open-shell energy is outside the closed-shell path")
    return _xl_energy(P, P, F, _sym(Hcore, doTriu))


def elec_energy_xl(D, P, F, Hcore):
    """sum D F - 0.5 sum (F - H) P per molecule (energy.py:76-88)."""
    return _xl_energy(D, P, F, _sym(Hcore, True))


def pair_nuclear_energy(Z, const, nmol, ni, nj, idxi, idxj, rij, rho0xi, rho0xj, alp, chi, gam, method="AM1",
                        parameters=None):
    """Core-core repulsion per pair (energy.py:91-134) for MNDO / AM1 / PM3."""
    if method not in _METHODS:
        raise ValueError("Supported Method: MNDO, AM1, PM3")
    dev = nat.require_device(rij.device)
    natoms = int(Z.shape[0])
    par = torch.zeros(max(natoms, 1), nat.NPAR, dtype=torch.float64, device=dev)
    par[:natoms, 31] = Z.to(torch.float64)
    par[:natoms, 29] = torch.as_tensor(np.asarray(TORE), dtype=torch.float64, device=dev)[Z.long()]
    par[:natoms, 11] = parameters[0].to(torch.float64)
    if method != "MNDO":
        _, K, L, M = parameters
        for g in range(K.shape[1]):
            par[:natoms, 12 + 3 * g] = K[:, g]
            par[:natoms, 13 + 3 * g] = L[:, g]
            par[:natoms, 14 + 3 * g] = M[:, g]
    npairs = int(rij.shape[0])
    out = torch.empty(max(npairs, 1), dtype=torch.float64, device=dev)
    i32 = lambda t: t.to(torch.int32).contiguous()
    ii, jj = i32(idxi), i32(idxj)
    r = rij.to(torch.float64).contiguous()
    g = gam.to(torch.float64).contiguous()
    nat.call_raw("pyseqm_core_repulsion_pairs", npairs, _METHODS[method], nat.ptr(ii), nat.ptr(jj), nat.ptr(r),
                 nat.ptr(g), nat.ptr(par), nat.ptr(out), nat.stream_ptr(dev))
    return out[:npairs]


def total_energy(nmol, pair_molid, EnucAB, Eelec):
    """(Etot, Enuc) per molecule (energy.py:177-192)."""
    Enuc = torch.zeros((nmol,), dtype=EnucAB.dtype, device=EnucAB.device)
    Enuc.index_add_(0, pair_molid, EnucAB)
    return Eelec + Enuc, Enuc


def elec_energy_isolated_atom(const, Z, uss, upp, gss, gpp, gsp, gp2, hsp):
    """Isolated-atom electronic energy per atom (energy.py:8-23)."""
    return (uss * const.ussc[Z] + upp * const.uppc[Z] + gss * const.gssc[Z] + gpp * const.gppc[Z]
            + gsp * const.gspc[Z] + gp2 * const.gp2c[Z] + hsp * const.hspc[Z])


def heat_formation(const, nmol, atom_molid, Z, Etot, Eiso, flag=True):
    """(Hf, sum Eiso) per molecule (energy.py:195-216)."""
    Eiso_sum = torch.zeros_like(Etot)
    Eiso_sum.index_add_(0, atom_molid, Eiso)
    if flag:
        eheat_sum = torch.zeros_like(Etot)
        eheat_sum.index_add_(0, atom_molid, const.eheat[Z])
        return Etot - Eiso_sum + eheat_sum, Eiso_sum
    return Etot - Eiso_sum, Eiso_sum
