"""sym_eig_trunc: the FP64 diagonalisation density step in the reference's padded layout
(seqm/seqm_functions/diag.py:110-241) on the HIP kernels: Gershgorin-shifted padding
(pyseqm_pad_for_eigh), batched FP64 eigensolver, P = 2 C_occ C_occ^T (pyseqm_density_from_evecs).
Eigenvectors are returned as the reference does (columns of the packed-size matrix); their signs
within degenerate subspaces are solver dependent."""
import torch

from .. import _native as nat
from ..layout import PackedBatch

_CACHE = {}


def _batch(nheavyatom, nH, nocc, molsize, device):
    key = (tuple(torch.as_tensor(nheavyatom).reshape(-1).tolist()), tuple(torch.as_tensor(nH).reshape(-1).tolist()),
           tuple(torch.as_tensor(nocc).reshape(-1).tolist()), int(molsize), str(device))
    b = _CACHE.get(key)
    if b is None:
        if len(_CACHE) > 64:
            _CACHE.clear()
        b = _CACHE[key] = PackedBatch.from_counts(nheavyatom, nH, nocc, molsize, device)
    return b


def sym_eig_trunc(x, nheavyatom, nH, nocc, eig_only=False):
    if x.dim() == 4:
        raise NotImplementedError("the unrestricted (nmol, 2, n, n) branch is outside the closed-shell path")
    single = x.dim() == 2
    if single:
        x = x.unsqueeze(0)
    dev = nat.require_device(x.device)
    size = x.shape[-1]
    b = _batch(nheavyatom, nH, nocc, size // 4, dev)
    F = b.from_padded(x.to(torch.float64))
    idx = torch.arange(b.nmol, dtype=torch.int32, device=dev)
    nmax = b.nmax
    red = torch.empty(b.nmol * nat.MAX_RED_BLOCKS * 16, dtype=torch.float64, device=dev)
    stream = nat.stream_ptr(dev)
    Fp = torch.empty(b.nmol, nmax, nmax, dtype=torch.float64, device=dev)
    nat.call("pyseqm_pad_for_eigh", b.L, nat.ptr(F), nat.ptr(idx), b.nmol, nmax, nat.ptr(Fp), nat.ptr(red),
             nat.red_blocks(nmax), stream)
    e0, v = torch.linalg.eigh(Fp)
    v = v.contiguous()
    e = torch.zeros(b.nmol, size, dtype=torch.float64, device=dev)
    mask = torch.arange(nmax, device=dev)[None, :] < torch.as_tensor(b.norb, device=dev)[:, None]
    e[:, :nmax] = torch.where(mask, e0, torch.zeros_like(e0))
    if eig_only:
        return (e[0], v[0]) if single else (e, v)
    P = torch.empty(b.total, dtype=torch.float64, device=dev)
    nat.call("pyseqm_density_from_evecs", b.L, nat.ptr(v), nmax, nat.ptr(idx), b.nmol, nat.ptr(P), stream)
    Pp = b.to_padded(P)
    if single:
        return e[0], Pp[0], v[0]
    return e, Pp, v
