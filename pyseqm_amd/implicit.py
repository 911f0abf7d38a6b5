"""SCF implicit backward on the device: gradients of density-dependent outputs with respect to the
learned per-atom parameters through the SCF fixed point (scf_loop.py:1779-1993, scf_backward = 1).

At the converged density P* = D(F(theta, P*)) of a closed-shell SCF (D: the aufbau density of the
nocc lowest eigenpairs), an upstream gradient gbar = dL/dP* gives

    u = gbar + G(R(u))                          (the reference's affine_eq, solved by Anderson steps)
    dL/dtheta = d/dtheta sum R(u) F(theta; P*)  (agrad(Pout, theta, grad_outputs=u))

with G the two-electron part of the Fock build (self-adjoint in the Frobenius product) and R the
first-order density response of D to a Fock perturbation,
R(X) = C [M o (C^T X C)] C^T, M_ia = M_ai = 2 / (e_i - e_a) for occupied i and virtual a
(self-adjoint as well).  Every piece is a device kernel: the eigenpairs (pyseqm_eigh_vectors; the
batched library eigensolver for blocks above its 139-orbital all-eigenpairs limit),
the eigenbasis products (pyseqm_block_gemm), M (pyseqm_density_response_elem), G (pyseqm_fock
with a zero core Hamiltonian) and the parameter contraction (pyseqm_fock_param_vjp).  The
Anderson mixing of the adjoint iterates runs as per-molecule torch reductions on the packed
blocks.
"""
import numpy as np
import torch

from . import _native as nat
from .engine import eigh_wg_ok


class ResponseWorkspace:
    """Device buffers of the density response and the adjoint iteration for one packed batch."""

    def __init__(self, eng):
        b = eng.b
        dev = eng.device
        norb, nocc = np.asarray(b.norb), np.asarray(b.nocc)
        if ((nocc < 1) | (nocc >= norb)).any():
            raise ValueError("the density response needs 1 <= nocc < norb in every molecule")
        f64 = dict(dtype=torch.float64, device=dev)
        total = max(b.total, 1)
        self.nmol, self.nmax = b.nmol, int(b.nmax)
        self.idx = torch.arange(b.nmol, dtype=torch.int32, device=dev)
        self.orb_off = torch.as_tensor(np.concatenate([[0], np.cumsum(norb)[:-1]]).astype(np.int64), device=dev)
        self.evals = torch.zeros(max(int(norb.sum()), 1), **f64)
        self.Q = torch.zeros(total, **f64)
        wg = norb[eigh_wg_ok(norb, norb)]
        self.eig_ws = torch.zeros(max(4 * b.nmol * int(wg.max()) ** 2 if len(wg) else 1, 1), **f64)
        self.fail = torch.zeros(1, dtype=torch.int32, device=dev)
        self.Hzero = torch.zeros(total, **f64)
        self.T1 = torch.zeros(total, **f64)
        self.T2 = torch.zeros(total, **f64)
        self.eelec = torch.zeros(b.nmol, **f64)
        n2 = norb.astype(np.int64) ** 2
        self.elem_mol = torch.as_tensor(np.repeat(np.arange(b.nmol), n2), device=dev)
        # per packed element (i, j) of molecule m: orb_off[m] + j (its column's eigenvalue slot); per
        # eigenvalue slot: (molecule, index) in the reference's padded orbital-energy layout
        off = np.concatenate([[0], np.cumsum(norb)[:-1]]).astype(np.int64)
        self.elem_col = torch.as_tensor(np.concatenate([np.tile(np.arange(n), n) + o for n, o in zip(norb, off)])
                                        if b.nmol else np.zeros(0, np.int64), device=dev)
        self.orb_mol = torch.as_tensor(np.repeat(np.arange(b.nmol), norb), device=dev)
        self.orb_k = torch.as_tensor(np.concatenate([np.arange(n) for n in norb]) if b.nmol
                                     else np.zeros(0, np.int64), device=dev)


def eigenpairs(eng, ws, F):
    ws.fail.zero_()
    eng.eigenpairs(F, ws.orb_off, ws.evals, ws.Q, ws.eig_ws, ws.fail)
    if int(ws.fail.item()):
        raise RuntimeError(f"eigensolver lost {int(ws.fail.item())} eigenvector(s) to a cluster's span: "
                           "the density response of the SCF backward is not defined for these blocks")


def _gemm(eng, ws, ta, tb, A, B, C):
    nat.call("pyseqm_block_gemm", eng.b.L, nat.ptr(ws.idx), eng.b.nmol, ws.nmax, int(ta), int(tb), nat.ptr(A),
             nat.ptr(B), nat.ptr(C), eng.stream)


def response(eng, ws, X, out):
    """out = R(X) = Q [M o (Q^T X Q)] Q^T for symmetric packed X."""
    _gemm(eng, ws, 0, 0, X, ws.Q, ws.T1)
    _gemm(eng, ws, 1, 0, ws.Q, ws.T1, ws.T2)
    nat.call("pyseqm_density_response_elem", eng.b.L, nat.ptr(ws.idx), eng.b.nmol, ws.nmax, nat.ptr(ws.orb_off),
             nat.ptr(ws.evals), nat.ptr(ws.T2), eng.stream)
    _gemm(eng, ws, 0, 0, ws.Q, ws.T2, ws.T1)
    _gemm(eng, ws, 0, 1, ws.T1, ws.Q, out)
    return out


def two_electron(eng, ws, X, out, atom_par, w):
    """out = G(X): the Fock build of X with a zero core Hamiltonian (integrals w = (w_xx, w_xh, w_hh))."""
    nat.call("pyseqm_fock", eng.b.L, nat.ptr(atom_par), nat.ptr(w[0]), nat.ptr(w[1]), nat.ptr(w[2]), nat.ptr(X),
             nat.ptr(ws.Hzero), nat.ptr(out), nat.ptr(eng.part), nat.ptr(eng.tile_v), nat.ptr(eng.atom_v),
             nat.ptr(ws.eelec), None, eng.stream)
    return out


def adjoint(eng, ws, gbar, atom_par, w, eps, max_iter=300, depth=6):
    """Solve u = gbar + G(R(u)) by Anderson-accelerated fixed-point steps (type II, per-molecule
    coefficients from the last `depth` residual differences).  Returns u and a per-molecule flag of
    molecules whose residual max |g(u) - u| stayed above eps."""
    nmol = ws.nmol
    em = ws.elem_mol
    seg_max = lambda x: torch.zeros(nmol, dtype=x.dtype, device=x.device).scatter_reduce_(0, em, x.abs(), "amax")
    dots = lambda x, y: torch.zeros(nmol, dtype=x.dtype, device=x.device).index_add_(0, em, x * y)
    Rv = torch.empty_like(gbar)
    Gv = torch.empty_like(gbar)
    u = gbar.clone()
    dF, dG = [], []
    f_prev = g_prev = None
    err = None
    for _ in range(max_iter):
        gu = gbar + two_electron(eng, ws, response(eng, ws, u, Rv), Gv, atom_par, w)
        f = gu - u
        err = seg_max(f)
        if bool((err <= eps).all()):
            return gu, err > eps
        if f_prev is not None:
            dF.append(f - f_prev)
            dG.append(gu - g_prev)
            if len(dF) > depth:
                dF.pop(0)
                dG.pop(0)
        f_prev, g_prev = f, gu
        if dF:
            k = len(dF)
            A = torch.stack([torch.stack([dots(dF[i], dF[j]) for j in range(k)], 1) for i in range(k)], 1)
            rhs = torch.stack([dots(dF[i], f) for i in range(k)], 1)
            reg = 1e-12 * A.diagonal(dim1=1, dim2=2).sum(1) + 1e-280  # stays a normal number
            A = A + reg[:, None, None] * torch.eye(k, dtype=A.dtype, device=A.device)
            gam = torch.linalg.solve(A, rhs[..., None])[..., 0]
            u = gu.clone()
            for i in range(k):
                u -= gam[:, i][em] * dG[i]
        else:
            u = gu
    return u, err > eps


def eigenvalue_source(eng, ws, ebar):
    """Y_e = sum_k ebar_k c_k c_k^T (packed) for an upstream gradient ebar on the orbital energies
    (one value per eigenvalue slot orb_off[m] + k): sum_k ebar_k de_k = sum Y_e dF."""
    T = ws.Q * ebar[ws.elem_col]
    Ye = torch.empty_like(T)
    _gemm(eng, ws, 0, 1, T, ws.Q, Ye)
    return Ye


def implicit_param_grad(eng, P, F, gbar, atom_par, coords, w, eps, hrel=1.0e-4, epad_grad=None):
    """dL/dtheta (natoms, 24) for the upstream gradient gbar = dL/dP* (packed, symmetrised) and,
    optionally, epad_grad = dL/de on the orbital energies in the reference's padded layout
    (nmol, 4 molsize); returns also the per-molecule flag of a backward that did not reach eps.
    The orbital energies e_k = c_k^T F(theta, P*) c_k add Y_e = sum_k ebar_k c_k c_k^T to the
    parameter contraction and G(Y_e) to the adjoint's source (F depends on P* through G)."""
    ws = ResponseWorkspace(eng)
    eigenpairs(eng, ws, F)
    Ye = None
    if epad_grad is not None:
        ebar = epad_grad.to(torch.float64)[ws.orb_mol, ws.orb_k].contiguous()
        Ye = eigenvalue_source(eng, ws, ebar)
        gbar = gbar + two_electron(eng, ws, Ye, torch.empty_like(Ye), atom_par, w)
    u, bad = adjoint(eng, ws, gbar, atom_par, w, eps)
    Y = response(eng, ws, u, torch.empty_like(u))
    if Ye is not None:
        Y = Y + Ye
    b = eng.b
    out = torch.zeros(max(b.natoms, 1), 2, 24, dtype=torch.float64, device=eng.device)
    rec = torch.empty(max(int(nat.lib().pyseqm_param_grad_workspace(nat.ctypes.byref(b.L))), 1),
                      dtype=torch.float64, device=eng.device)
    nat.call("pyseqm_fock_param_vjp", b.L, nat.METHODS[eng.method], nat.ptr(coords), nat.ptr(atom_par), nat.ptr(P),
             nat.ptr(Y), float(hrel), nat.ptr(rec), nat.ptr(eng.part), nat.ptr(out), eng.stream)
    return out[:b.natoms, 0], bad


class SCFImplicit(torch.autograd.Function):
    """The converged density (padded, as scf_loop returns it) and, when given, the orbital energies
    (padded, eig=True) as functions of the learned per-atom parameters, with the implicit backward
    above.  The engine's parameters, integrals and
    coordinates at forward time are kept, so a later re-parametrisation does not leak into the
    backward.  Molecules whose forward SCF or backward iteration did not converge get zero
    gradients, as in the reference (scf_loop.py:1950-1975)."""

    @staticmethod
    def forward(ctx, eng, P, F, eps, notconverged, names, Ppad, epad, *params):
        ctx.eng, ctx.names, ctx.eps = eng, names, eps
        ctx.state = (P.detach().clone(), F.detach().clone(), eng.atom_par.clone(), eng.coords.clone(),
                     (eng.w_xx.clone(), eng.w_xh.clone(), eng.w_hh.clone()),
                     torch.as_tensor(notconverged, device=P.device).bool().clone())
        ctx.meta = [(p.shape, p.dtype, p.device) for p in params]
        ctx.has_e = epad is not None
        return (Ppad.clone(), epad.clone()) if ctx.has_e else Ppad.clone()

    @staticmethod
    def backward(ctx, gP, ge=None):
        import warnings
        eng = ctx.eng
        b = eng.b
        P, F, atom_par, coords, w, nc = ctx.state
        gs = 0.5 * (gP + gP.transpose(1, 2))
        gbar = b.from_padded(gs.to(torch.float64)).contiguous()
        grad, bad = implicit_param_grad(eng, P, F, gbar, atom_par, coords, w, ctx.eps,
                                        epad_grad=ge if ctx.has_e else None)
        drop = nc | bad
        if bool(bad.any()):
            warnings.warn("SCF backward for %d/%d molecules did not converge" % (int(bad.sum()), b.nmol))
        if bool(drop.any()):
            grad = grad * (~drop)[torch.as_tensor(b.atom_mol, device=grad.device).long()][:, None]
        from .seqm_functions.parameters import COLUMNS
        grads = [grad[:, COLUMNS.index(n)].reshape(shape).to(dtype=dtype, device=dev)
                 for n, (shape, dtype, dev) in zip(ctx.names, ctx.meta)]
        return (None, None, None, None, None, None, None, None, *grads)
