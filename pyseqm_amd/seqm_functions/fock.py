"""fock(...): F = M + G(P) with the reference's signature and padded layouts
(seqm/seqm_functions/fock.py:132-347) on the HIP Fock-build kernel.

The caller's inputs are brought into the packed layout the kernel reads: the atoms from
maskd / Z, the core Hamiltonian from the block array M (upper triangle, hcore.py:155-179),
the two-center integrals w (npairs, 10, 10) into the per-tile compact layout (pairs missing
from a cut-off pair list get zero integrals, i.e. no interaction, as in the reference), and the
one-center parameters gss, gpp, gsp, gp2, hsp per atom."""
import numpy as np
import torch

from .. import _native as nat
from ..layout import PackedBatch

_CACHE = {}


def species_from_masks(nmol, molsize, maskd, Z):
    """Padded species (nmol, molsize) from the diagonal-block positions maskd = mol m^2 + i (m + 1)."""
    sp = torch.zeros(nmol * molsize, dtype=torch.int64, device=Z.device)
    sp[torch.div(maskd, molsize, rounding_mode="floor")] = Z.long()
    return sp.reshape(nmol, molsize)


def batch_for(species, method, device):
    sp = species.detach().cpu().numpy()
    key = (sp.tobytes(), sp.shape, method, str(device))
    b = _CACHE.get(key)
    if b is None:
        if len(_CACHE) > 32:
            _CACHE.clear()
        b = _CACHE[key] = PackedBatch(sp, method, device)
    return b


def compact_w(b, w, idxi, idxj):
    """Dense (npairs, 10, 10) integrals in the reference pair order -> per-tile compact buffers."""
    dev = w.device
    # position of each given pair in the full (mol, i < j) list the tiles index
    a0 = torch.as_tensor(b.mol_atom_off, device=dev)
    mol = torch.as_tensor(b.atom_mol, device=dev).long()[idxi.long()]
    li = idxi.long() - a0[mol]
    lj = idxj.long() - a0[mol]
    nat_m = torch.as_tensor(b.nat, device=dev)[mol]
    full = torch.as_tensor(b.pair_base, device=dev)[mol] + li * nat_m - li * (li + 1) // 2 + (lj - li - 1)
    wf = torch.zeros(max(b.npairs, 1), 10, 10, dtype=torch.float64, device=dev)
    wf[full] = w.to(torch.float64)
    out = []
    for kind in range(3):
        buf = torch.zeros(max(b.w_size[kind], 1), dtype=torch.float64, device=dev)
        ref, idx = b.w_gather_index(kind, dev)
        if ref is not None:
            if kind == 0:
                buf[idx] = wf[ref].reshape(-1, 100)
            elif kind == 1:
                buf[idx] = wf[ref, :, 0]
            else:
                buf[idx] = wf[ref, 0, 0].unsqueeze(1)
        out.append(buf)
    return out


def blocks_to_matrix(M, nmol, molsize):
    """Reference block array (nmol m^2, 4, 4) -> symmetric (nmol, 4m, 4m) from its upper triangle."""
    Hp = M.reshape(nmol, molsize, molsize, 4, 4).transpose(2, 3).reshape(nmol, 4 * molsize, 4 * molsize)
    return Hp.triu() + Hp.triu(1).transpose(1, 2)


def matrix_to_blocks(Hp, nmol, molsize):
    """Symmetric (nmol, 4m, 4m) -> reference block array holding its upper triangle."""
    return Hp.triu().reshape(nmol, molsize, 4, molsize, 4).transpose(2, 3).reshape(nmol * molsize * molsize, 4, 4)


def fock(nmol, molsize, P0, M, maskd, mask, idxi, idxj, w, W, gss, gpp, gsp, gp2, hsp, themethod, zetas, zetap,
         zetad, Z, F0SD, G2SD):
    if P0.dim() == 4:
        raise NotImplementedError("the unrestricted Fock build is outside the closed-shell path")
    dev = nat.require_device(P0.device)
    sp = species_from_masks(nmol, molsize, maskd, Z)
    b = batch_for(sp, themethod, dev)
    par = torch.zeros(max(b.natoms, 1), nat.NPAR, dtype=torch.float64, device=dev)
    for col, v in ((6, gss), (7, gsp), (8, gpp), (9, gp2), (10, hsp)):
        par[:b.natoms, col] = v.to(torch.float64)
    w_xx, w_xh, w_hh = compact_w(b, w, idxi, idxj)
    H = b.from_padded(blocks_to_matrix(M.to(torch.float64), nmol, molsize))
    P = b.from_padded(P0.to(torch.float64))
    F = torch.empty_like(P)
    nt = max(b.ntiles, 1)
    part = torch.empty(nt * 160, dtype=torch.float64, device=dev)
    tile_v = torch.empty(nt, dtype=torch.float64, device=dev)
    atom_v = torch.empty(max(b.natoms, 1), dtype=torch.float64, device=dev)
    eelec = torch.empty(max(b.nmol, 1), dtype=torch.float64, device=dev)
    nat.call("pyseqm_fock", b.L, nat.ptr(par), nat.ptr(w_xx), nat.ptr(w_xh), nat.ptr(w_hh), nat.ptr(P), nat.ptr(H),
             nat.ptr(F), nat.ptr(part), nat.ptr(tile_v), nat.ptr(atom_v), nat.ptr(eelec), None,
             nat.stream_ptr(dev))
    return b.to_padded(F).to(P0.dtype)
