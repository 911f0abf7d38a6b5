"""Thermal density and the Krylov-subspace (KSA) XL-BOMD kernel on the device: the max_rank branch
of EnergyXL (seqm/dynamics/xlbomd.py:201-339) with Fermi_Q (seqm/seqm_functions/fermi_q.py:8-72),
G (seqm/seqm_functions/G_XL_LR.py:7-908) and Canon_DM_PRT (seqm/seqm_functions/canon_dm_prt.py:6-39).

Per step, on the packed batch:

    Q, e = eigenpairs of F(P)                             (pyseqm_eigh_vectors)
    mu, f, D = 2 Q diag(f) Q^T, S                          (pyseqm_fermi_density)
    Krylov loop k = 0 .. rank-1 (while the fit's largest residual exceeds err_threshold):
        V_k = W_{k-1} (W_{-1} = D - P), Gram-Schmidt against V_j, normalised   (block dot / axpy)
        FO1 = G(V_k)                                     (the Fock kernels with H = 0)
        W_k = Q canon(Q^T FO1 Q) Q^T - V_k               (block GEMMs, pyseqm_canon_prt)
        alpha = argmin |sum_r alpha_r W_r - (D - P)|     (Frobenius products, pyseqm_small_solve)
    dP2dt2 = -sum_r alpha_r V_r

Every array stays on the device; with err_threshold <= 0 (the reference tests' setting) the loop
length is fixed and nothing is read back.
"""
import numpy as np
import torch

from . import _native as nat
from .engine import eigh_wg_ok

KB_EV = 8.61739e-5      # Boltzmann constant, eV/K (xlbomd.py:214)
CANON_DM_PRT_ITER = 10  # recursion steps of Canon_DM_PRT in EnergyXL (xlbomd.py:55)
K0 = 1.0                # xlbomd.py:248


class KSAWorkspace:
    """Device buffers of the thermal density and the rank-m Krylov fit for one packed batch."""

    def __init__(self, eng, rank):
        b = eng.b
        dev = eng.device
        self.R = int(nat.lib().pyseqm_ksa_max_rank())
        if not 1 <= int(rank) <= self.R:
            raise NotImplementedError(f"max_rank = {rank}: the device Krylov fit takes 1 .. {self.R}")
        norb, nocc = np.asarray(b.norb), np.asarray(b.nocc)
        if int(norb.max()) > 256:
            raise NotImplementedError("the thermal density and its perturbation recursion take blocks of up to 256 "
                                      f"orbitals (largest here: {int(norb.max())})")
        if ((nocc < 1) | (nocc >= norb)).any():
            raise ValueError("the Fermi density needs 1 <= nocc < norb in every molecule (HOMO and LUMO)")
        f64 = dict(dtype=torch.float64, device=dev)
        total = max(b.total, 1)
        self.rank = int(rank)
        self.orb_off = torch.as_tensor(np.concatenate([[0], np.cumsum(norb)[:-1]]).astype(np.int64), device=dev)
        nsum = max(int(norb.sum()), 1)
        self.evals = torch.zeros(nsum, **f64)
        self.occ = torch.zeros(nsum, **f64)
        self.mu = torch.zeros(b.nmol, **f64)
        self.S = torch.zeros(b.nmol, **f64)
        self.Q = torch.zeros(total, **f64)
        self.work = torch.zeros(int(nat.lib().pyseqm_fermi_workspace(b.nmol)), dtype=torch.uint8, device=dev)
        wg = norb[eigh_wg_ok(norb, norb)]
        self.eig_ws = torch.zeros(max(4 * b.nmol * int(wg.max()) ** 2 if len(wg) else 1, 1), **f64)
        self.fail = torch.zeros(1, dtype=torch.int32, device=dev)
        self.Hzero = torch.zeros(total, **f64)
        self.G = torch.zeros(total, **f64)
        self.T1 = torch.zeros(total, **f64)
        self.Y = torch.zeros(total, **f64)
        self.dDS = torch.zeros(total, **f64)
        self.res = torch.zeros(total, **f64)
        self.V = torch.zeros(self.rank, total, **f64)
        self.W = torch.zeros(self.rank, total, **f64)
        self.dot = torch.zeros(b.nmol, **f64)
        self.nrm = torch.zeros(b.nmol, **f64)
        self.O = torch.zeros(b.nmol, self.R, self.R, **f64)
        self.rhs = torch.zeros(b.nmol, self.R, **f64)
        self.alpha = torch.zeros(b.nmol, self.R, **f64)
        self.eelec = torch.zeros(b.nmol, **f64)
        self.idx = torch.arange(b.nmol, dtype=torch.int32, device=dev)
        self.last_error = None
        self.last_rank = 0


def _dot(eng, ks, A, B, out):
    nat.call("pyseqm_block_dot", eng.b.L, nat.ptr(ks.idx), eng.b.nmol, nat.ptr(A), nat.ptr(B), nat.ptr(out),
             eng.stream)


def _axpy(eng, ks, mode, s, c, cstride, X, Y):
    b = eng.b
    nat.call("pyseqm_block_axpy", b.L, nat.ptr(ks.idx), b.nmol, int(b.nmax), int(mode), float(s),
             nat.ptr(c), int(cstride), nat.ptr(X), nat.ptr(Y), eng.stream)


def _gemm(eng, ks, ta, tb, A, B, C):
    b = eng.b
    nat.call("pyseqm_block_gemm", b.L, nat.ptr(ks.idx), b.nmol, int(b.nmax), int(ta), int(tb), nat.ptr(A),
             nat.ptr(B), nat.ptr(C), eng.stream)


def fermi_density(eng, ks, F, T, D, kB=KB_EV):
    """Fermi_Q of the Fock matrices F into D (packed); eigenpairs, occupations, mu and the entropy
    stay in the workspace (ks.Q, ks.evals, ks.occ, ks.mu, ks.S)."""
    b = eng.b
    nmax = int(b.nmax)
    eng.eigenpairs(F, ks.orb_off, ks.evals, ks.Q, ks.eig_ws, ks.fail)
    nat.call("pyseqm_fermi_density", b.L, nat.ptr(ks.idx), b.nmol, nmax, nat.ptr(ks.orb_off), nat.ptr(ks.evals),
             nat.ptr(ks.Q), float(T), float(kB), nat.ptr(ks.occ), nat.ptr(ks.mu), nat.ptr(D), nat.ptr(ks.S),
             nat.ptr(ks.work), eng.stream)


def g_matrix(eng, ks, V, out):
    """G(V) (G_XL_LR.py:7-908): the two-electron Fock matrix of a density V, no core Hamiltonian."""
    b = eng.b
    nat.call("pyseqm_fock", b.L, nat.ptr(eng.atom_par), nat.ptr(eng.w_xx), nat.ptr(eng.w_xh), nat.ptr(eng.w_hh),
             nat.ptr(V), nat.ptr(ks.Hzero), nat.ptr(out), nat.ptr(eng.part), nat.ptr(eng.tile_v),
             nat.ptr(eng.atom_v), nat.ptr(ks.eelec), None, eng.stream)


def canon_response(eng, ks, FO1, out, T, kB=KB_EV, steps=CANON_DM_PRT_ITER):
    """Canon_DM_PRT: out = Q canon(Q^T FO1 Q) Q^T with the trace-keeping mu shift."""
    b = eng.b
    _gemm(eng, ks, 0, 0, FO1, ks.Q, ks.T1)       # FO1 Q
    _gemm(eng, ks, 1, 0, ks.Q, ks.T1, ks.Y)      # Q^T (FO1 Q)
    nat.call("pyseqm_canon_prt", b.L, nat.ptr(ks.idx), b.nmol, int(b.nmax), nat.ptr(ks.orb_off), nat.ptr(ks.evals),
             nat.ptr(ks.mu), float(T), float(kB), int(steps), nat.ptr(ks.Y), eng.stream)
    _gemm(eng, ks, 0, 0, ks.Q, ks.Y, ks.T1)      # Q P1
    _gemm(eng, ks, 0, 1, ks.T1, ks.Q, out)       # (Q P1) Q^T


def _fit(eng, ks, r):
    """alpha = argmin over the first r responses: O alpha = rhs, O_ij = <W_i, W_j>, rhs_i = <W_i, D - P>."""
    for i in range(r):
        for j in range(r):
            _dot(eng, ks, ks.W[i], ks.W[j], ks.dot)
            ks.O[:, i, j].copy_(ks.dot)
        _dot(eng, ks, ks.W[i], ks.dDS, ks.dot)
        ks.rhs[:, i].copy_(ks.dot)
    nat.call_raw("pyseqm_small_solve", nat.ptr(ks.idx), eng.b.nmol, int(r), nat.ptr(ks.O), nat.ptr(ks.rhs),
                 nat.ptr(ks.alpha), eng.stream)


def rank_kernel(eng, ks, D, P, T, err_threshold, dP2dt2, kB=KB_EV):
    """The rank-m Krylov approximation dP2dt2 = -sum_r alpha_r V_r of EnergyXL (xlbomd.py:243-339)."""
    R = ks.R
    ks.dDS.copy_(D)
    _axpy(eng, ks, 4, -K0, None, 0, P, ks.dDS)   # K0 (D - P), K0 = 1
    dW = ks.dDS
    k = -1
    err = None
    if err_threshold > 0.0:
        _dot(eng, ks, ks.dDS, ks.dDS, ks.nrm)
    while k < ks.rank - 1:
        k += 1
        Vk = ks.V[k]
        Vk.copy_(dW)
        for j in range(k):  # modified Gram-Schmidt in the Frobenius product
            _dot(eng, ks, Vk, ks.V[j], ks.dot)
            _axpy(eng, ks, 0, -1.0, ks.dot, 1, ks.V[j], Vk)
        _dot(eng, ks, Vk, Vk, ks.dot)
        _axpy(eng, ks, 1, 1.0, ks.dot, 1, None, Vk)
        g_matrix(eng, ks, Vk, ks.G)
        canon_response(eng, ks, ks.G, ks.W[k], T, kB)
        _axpy(eng, ks, 2, K0, None, 0, Vk, ks.W[k])  # W_k = K0 (P1 - V_k)
        dW = ks.W[k]
        _fit(eng, ks, k + 1)
        if err_threshold > 0.0:
            _axpy(eng, ks, 3, 1.0, ks.alpha[:, 0], R, ks.W[0], ks.res)
            for r in range(1, k + 1):
                _axpy(eng, ks, 0, 1.0, ks.alpha[:, r], R, ks.W[r], ks.res)
            _axpy(eng, ks, 4, -1.0, None, 0, ks.dDS, ks.res)
            _dot(eng, ks, ks.res, ks.res, ks.dot)
            err = torch.sqrt(ks.dot / ks.nrm)
            if float(err.max()) <= err_threshold:  # the reference's loop test (host read only here)
                break
    # dP2dt2 = -sum_r alpha_r V_r with the last fit
    _axpy(eng, ks, 3, -1.0, ks.alpha[:, 0], R, ks.V[0], dP2dt2)
    for r in range(1, k + 1):
        _axpy(eng, ks, 0, -1.0, ks.alpha[:, r], R, ks.V[r], dP2dt2)
    ks.last_error = err
    ks.last_rank = k + 1
    return dP2dt2
