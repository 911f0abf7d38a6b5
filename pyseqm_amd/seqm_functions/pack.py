"""pack / unpack between the reference's padded orbital layout (4 rows per atom, hydrogens use the
first row of their block) and the compact per-molecule layout (seqm/seqm_functions/pack.py:8-96).

One gather/scatter per batch on the device (no per-molecule loop); molecules smaller than the
largest are zero-padded at the end, as in the reference."""
import torch


def _rows(nHeavy, nHydro, device):
    nho = 4 * torch.as_tensor(nHeavy, device=device).reshape(-1).long()
    nH = torch.as_tensor(nHydro, device=device).reshape(-1).long()
    norb = nho + nH
    N = int(norb.max().item()) if norb.numel() else 0
    k = torch.arange(N, device=device)[None, :]
    rows = torch.where(k < nho[:, None], k, nho[:, None] + 4 * (k - nho[:, None]))
    valid = k < norb[:, None]
    return torch.where(valid, rows, torch.zeros_like(rows)), valid, N


def pack(x, nHeavy, nHydro):
    single = x.dim() == 2
    if single:
        x = x.unsqueeze(0)
    elif x.dim() == 4:
        x = x.flatten(0, 1)
        nHeavy = torch.as_tensor(nHeavy).repeat_interleave(2)
        nHydro = torch.as_tensor(nHydro).repeat_interleave(2)
    rows, valid, N = _rows(nHeavy, nHydro, x.device)
    b = torch.arange(x.shape[0], device=x.device)[:, None, None]
    x0 = x[b, rows[:, :, None], rows[:, None, :]]
    x0 = x0 * (valid[:, :, None] & valid[:, None, :]).to(x0.dtype)
    return x0[0] if single else x0


def unpack(x0, nHeavy, nHydro, size):
    single = x0.dim() == 2
    if single:
        x0 = x0.unsqueeze(0)
    rows, valid, N = _rows(nHeavy, nHydro, x0.device)
    x = x0.new_zeros(x0.shape[0], size, size)
    mask = valid[:, :, None] & valid[:, None, :]
    b = torch.arange(x0.shape[0], device=x0.device)[:, None, None].expand_as(mask)
    r = rows[:, :, None].expand_as(mask)
    c = rows[:, None, :].expand_as(mask)
    x[b[mask], r[mask], c[mask]] = x0[:, :N, :N][mask]
    return x[0] if single else x
