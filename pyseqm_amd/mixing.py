"""Adaptive density mixing (scf_converger [1]) on the packed layout.

Restates adaptive_mix / compute_fac (seqm/seqm_functions/scf_loop.py:350-420): every third
iteration the off-diagonal density is extrapolated with FAC = sqrt(|dD|^2 / |d2D|^2); the
diagonal is damped (|step| <= 0.05 after iteration 4), clamped to [0, 2] and renormalised to the
new trace in at most 20 passes.  Element-wise work runs as torch device ops on the flat packed
buffers; the matrices never leave the device.
"""
import numpy as np
import torch


class AdaptiveMixer:
    def __init__(self, eng):
        b = eng.b
        dev = eng.device
        didx, dmol = b.diag_index()
        self.didx = torch.as_tensor(didx, device=dev)
        self.dmol = torch.as_tensor(dmol, device=dev)
        mat_mol = np.repeat(np.arange(b.nmol), b.norb.astype(np.int64) ** 2)
        self.mat_mol = torch.as_tensor(mat_mol, device=dev)
        self.old2 = torch.zeros(len(didx), dtype=torch.float64, device=dev)
        self.nmol = b.nmol
        self.dev = dev

    def _segsum(self, x, sel_mol):
        out = torch.zeros(self.nmol, dtype=x.dtype, device=self.dev)
        return out.index_add_(0, sel_mol, x)

    def step(self, k, P, Pnew, Pold, idx, idx_host):
        act_mol = torch.zeros(self.nmol, dtype=torch.bool, device=self.dev)
        act_mol[idx.long()] = True
        osel = act_mol[self.dmol]                 # active orbitals
        di_idx = self.didx[osel]
        dm = self.dmol[osel]
        dprev = P[di_idx]
        dcur = Pnew[di_idx]
        old2 = self.old2[osel]
        third = k % 3 == 0
        damp = 0.05 if k > 4 else 1.0e10
        if third:
            d1 = dcur - dprev
            d2 = dcur - 2.0 * dprev + old2
            num = self._segsum(d1 * d1, dm)
            den = self._segsum(d2 * d2, dm)
            valid = (den > 0) & (num < 100.0 * den)
            fac = torch.where(valid, torch.sqrt(torch.where(valid, num / torch.where(den > 0, den, 1.0), 0.0)),
                              torch.zeros_like(num))
        else:
            fac = torch.zeros(self.nmol, dtype=torch.float64, device=self.dev)
        esel = act_mol[self.mat_mol]
        Pold[esel] = P[esel]
        if third:
            f = fac[self.mat_mol[esel]]
            P[esel] = (1.0 + f) * Pnew[esel] - f * P[esel]
        else:
            P[esel] = Pnew[esel]
        delta = dcur - dprev
        cap = delta.abs() > damp
        dnew = torch.where(cap, dprev + torch.sign(delta) * damp, dcur + fac[dm] * delta).clamp(0.0, 2.0)
        sum0 = self._segsum(dcur, dm)
        di = dnew
        actv = act_mol
        for _ in range(20):
            sum2 = self._segsum(di, dm)
            large = sum2 > 1.0e-3
            sum3 = torch.where(large, sum0 / torch.where(large, sum2, torch.ones_like(sum2)), torch.zeros_like(sum2))
            done = (~large) | ((sum3 - 1.0).abs() <= 1.0e-5)
            if bool(torch.all(done[actv])):
                break
            scaled = (di * sum3[dm]).clamp(min=0.0)
            full = scaled > 2.0
            di = torch.where(full, torch.full_like(scaled, 2.0), scaled)
            sum0 = sum0 - self._segsum(full.to(di.dtype), dm) * 2.0
        P[di_idx] = di
        o2 = self.old2.clone()
        o2[osel] = dprev
        self.old2 = o2
