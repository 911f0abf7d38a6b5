"""Device SCF engine: drives the gfx950 kernels of libpyseqm_hip.so on one packed batch.

Host code only schedules kernels and reads one byte per molecule per SCF iteration (the
convergence flags), like the reference's ``int(notconverged.sum())`` host sync
(seqm/seqm_functions/scf_loop.py:1101).  Every floating-point array lives on the device.
"""
import collections
import math
import warnings

import numpy as np
import torch

from . import _native as nat
from .layout import PackedBatch
from .seqm_functions.constants import EHEAT, GP2C, GPPC, GSPC, GSSC, HSPC, TORE, UPPC, USSC, ev_kcalpmol
from .seqm_functions.parameters import COLUMNS, param_table

MAX_ITER = 1000
NFOCK = 10
SP2_DIIS_TOL = 1.0e-2


class SCFResult:
    pass


# Batched FP64 eigensolves of blocks with n <= 64 take a far slower solver path than n >= 68
# (measured on MI355X with the ROCm 7 solver: 1024 blocks of 64 take 417 ms, of 68 take 8 ms), so
# a group of many small blocks is padded up to this size (on 1024 alkanes a floor of 32 or 48 made
# the FP64 SCF 2.3-2.9x slower, 96 or 128 3-33 % slower).
EIGH_MIN_N = 68
# the floor applies to batched solves of at least this many blocks (a handful of small blocks is
# solved unpadded; the slow small-block path only costs at batch scale)
EIGH_MIN_N_BATCH = 16
# largest block whose DIIS residual stays on the FP64 GEMM with the mixed-precision density (the
# bf16x3 commutator is faster per iteration, but on small blocks its coarser residual costs SCF
# iterations: C2 converged in 11 instead of 9)
MIXED_DIIS_MIN_N = 256


# Blocks up to this size (and whose LDS footprint fits, eigh_wg_ok) are diagonalised by the
# one-workgroup-per-molecule solver of eigh.hip, which picks its thread grid by the launch's
# largest block: up to 48, 96 and 192 orbitals (EIGH_WG_CLASSES).
EIGH_WG_MAX_N = 192
EIGH_WG_LDS_DOUBLES = 20480
EIGH_WG_CLASSES = (48, 96, 192)


def eigh_wg_lds(norb, nvec):
    """LDS doubles of the workgroup eigensolver for a launch whose largest block has norb orbitals
    and nvec vectors: the norb x nvec eigenvectors, 3 norb (diagonal, off-diagonal, eigenvalues),
    three buffers of the variant's 48 / 96 / 192 grid rows, 128 of scratch (eigh.hip lds_doubles)."""
    n = np.asarray(norb, dtype=np.int64)
    o = np.asarray(nvec, dtype=np.int64)
    slots = np.where(n <= 48, 48, np.where(n <= 96, 96, 192))
    return n * o + 3 * n + 3 * slots + 128


def eigh_wg_ok(norb, nocc):
    """Whether the workgroup eigensolver takes blocks of norb orbitals / nocc occupied orbitals:
    n <= 192 and its LDS footprint (eigh_wg_lds) fits 160 KiB (pyseqm_eigh_density_lds_ok)."""
    n = np.asarray(norb, dtype=np.int64)
    o = np.asarray(nocc, dtype=np.int64)
    return (n >= 1) & (n <= EIGH_WG_MAX_N) & (o <= n) & (eigh_wg_lds(n, o) <= EIGH_WG_LDS_DOUBLES)


def size_groups(norb, ratio=0.7, min_group=32, floor=EIGH_MIN_N):
    """Partition molecules (by position) into groups of similar orbital counts: sorted by size,
    a group takes molecules down to `ratio` of its largest; groups smaller than `min_group` join
    the next larger one; sizes below `floor` count as `floor` (they are padded to it).  Returns a
    list of index arrays.  ratio 0.7 was the fastest on 1024 alkanes of 14-182 orbitals (3 groups,
    374 ms per FP64 SCF against 394 at 0.8 / 5 groups, 392 at 0.6 / 2 groups, 496 with one group):
    each batched solve costs a fixed chain of small kernels, which outweighs some padding."""
    norb = np.maximum(np.asarray(norb), floor)
    order = np.argsort(-norb, kind="stable")
    groups, start = [], 0
    while start < len(order):
        top = norb[order[start]]
        end = start
        while end < len(order) and norb[order[end]] >= ratio * top:
            end += 1
        if groups and len(groups[-1]) < min_group:
            groups[-1] = np.concatenate([groups[-1], order[start:end]])
        else:
            groups.append(order[start:end])
        start = end
    if len(groups) > 1 and len(groups[-1]) < min_group:
        groups[-2] = np.concatenate([groups[-2], groups[-1]])
        groups.pop()
    return groups


class SCFEngine:
    """Pair integrals, Fock builds, density steps and the SCF loop for one packed batch."""

    def __init__(self, species, coordinates, method, device="cuda", charges=None, batch=None, param_dir=None):
        self.device = nat.require_device(device)
        nat.lib()
        self.method = method
        self.b = batch if batch is not None else PackedBatch(species, method, self.device, charges)
        b = self.b
        dev = self.device
        f64 = dict(dtype=torch.float64, device=dev)
        tab = param_table(method, param_dir)
        self.default_params = True  # atom_par holds the parameter file's values
        par = np.zeros((b.natoms, nat.NPAR))
        par[:, :len(COLUMNS)] = tab[b.Z]
        self.atom_par = torch.as_tensor(par, **f64)
        self.stream = nat.stream_ptr(dev)
        nat.call("pyseqm_atom_setup", b.L, nat.ptr(self.atom_par), self.stream)
        nt = b.ntiles
        self.w_xx = torch.empty(max(b.w_size[0], 1), **f64)
        self.w_xh = torch.empty(max(b.w_size[1], 1), **f64)
        self.w_hh = torch.empty(max(b.w_size[2], 1), **f64)
        self.H = torch.empty(max(b.total, 1), **f64)
        self.part = torch.empty(max(nt, 1) * 160, **f64)
        self.tile_v = torch.empty(max(nt, 1), **f64)
        self.atom_v = torch.empty(max(b.natoms, 1), **f64)
        self.enuc = torch.zeros(b.nmol, **f64)
        self.red = torch.zeros(max(b.nmol, 1) * nat.MAX_RED_BLOCKS * 16, **f64)  # reduction scratch
        self.eelec = torch.zeros(b.nmol, **f64)
        # isolated-atom energies and atomic heats of formation (geometry independent)
        Z = b.Z
        eiso = (tab[Z, 0] * np.take(USSC, Z) + tab[Z, 1] * np.take(UPPC, Z) + tab[Z, 6] * np.take(GSSC, Z)
                + tab[Z, 8] * np.take(GPPC, Z) + tab[Z, 7] * np.take(GSPC, Z) + tab[Z, 9] * np.take(GP2C, Z)
                + tab[Z, 10] * np.take(HSPC, Z))
        self.eiso_atom = eiso
        self.eiso = torch.as_tensor(np.bincount(b.atom_mol, eiso, minlength=b.nmol), **f64)
        self.eheat = torch.as_tensor(np.bincount(b.atom_mol, np.take(EHEAT, Z) / ev_kcalpmol, minlength=b.nmol),
                                     **f64)
        self.coords = None
        self.trace = None  # set to a list to collect per-iteration (dE, diis, dm, dm_max) metrics
        self.diis_log = None  # set to a list to collect (cF, slot, B matrices, coefficients, resets, errors)
        self.eigh_log = collections.deque(maxlen=256)
        # vectors the workgroup eigensolver lost to Gram-Schmidt (stays 0 unless its vectors degenerate)
        self.eigh_fail = torch.zeros(1, dtype=torch.int32, device=dev)  # (blocks, padded size) of the recent batched eigensolves
        if coordinates is not None:
            self.set_coordinates(coordinates)

    # per-atom parameters the sp methods zero or derive (Molecule.py:101-113, basics.py:791-811): a
    # value passed for them does not change the reference's result either
    DERIVED_PARAMS = ("beta", "zeta_d", "s_orb_exp_tail", "p_orb_exp_tail", "d_orb_exp_tail", "U_dd", "F0SD",
                      "G2SD", "rho_core")

    def set_parameters(self, params):
        """Override per-atom MOPAC parameters ({name: (natoms,) tensor} in the batch's atom order, the
        dict Pack_Parameters returns with learned values merged in, basics.py:406-448, 781-789) and
        re-derive the multipoles, isolated-atom energies and, if coordinates are set, the integrals.
        Names the reference ignores for MNDO/AM1/PM3 are ignored; a per-pair ``Kbeta`` scaling of the
        resonance integrals (hcore.py:166-171) is not on this path and raises."""
        b = self.b
        kb = params.get("Kbeta", None)
        if torch.is_tensor(kb):
            raise NotImplementedError("per-pair Kbeta resonance scaling is not supported by the device engine")
        for name, v in params.items():
            if name not in COLUMNS or v is None:
                continue
            v = torch.as_tensor(v, dtype=torch.float64, device=self.device).detach().reshape(-1)
            if v.numel() != b.natoms:
                raise ValueError(f"parameter {name!r}: {v.numel()} values for {b.natoms} atoms")
            self.atom_par[:, COLUMNS.index(name)] = v
        self.default_params = False
        nat.call("pyseqm_atom_setup", b.L, nat.ptr(self.atom_par), self.stream)
        Z = torch.as_tensor(b.Z, device=self.device).long()
        ap = self.atom_par
        tab = lambda t: torch.as_tensor(t, dtype=torch.float64, device=self.device)[Z]
        eiso = (ap[:, 0] * tab(USSC) + ap[:, 1] * tab(UPPC) + ap[:, 6] * tab(GSSC) + ap[:, 8] * tab(GPPC)
                + ap[:, 7] * tab(GSPC) + ap[:, 9] * tab(GP2C) + ap[:, 10] * tab(HSPC))
        self.eiso_atom = eiso.cpu().numpy()
        self.eiso = torch.zeros(b.nmol, dtype=torch.float64, device=self.device).index_add_(
            0, torch.as_tensor(b.atom_mol, device=self.device).long(), eiso)
        if self.coords is not None:
            self.set_coordinates(self.coords)

    # ------------------------------------------------------------- integrals
    def set_coordinates(self, coordinates, enuc_pair=False):
        """coordinates: (nmol, molsize, 3) Angstrom in the reference's padded layout, or (natoms, 3)."""
        b = self.b
        c = torch.as_tensor(coordinates, dtype=torch.float64, device=self.device)
        if c.dim() == 3:
            c = c.reshape(-1, 3)[torch.as_tensor(b.real_flat, device=self.device)]
        self.coords = c.contiguous()
        self.enuc_pair = torch.zeros(max(b.npairs, 1), dtype=torch.float64, device=self.device) if enuc_pair \
            else None
        nat.call("pyseqm_pair_integrals", b.L, nat.METHODS[self.method], nat.ptr(self.coords),
                 nat.ptr(self.atom_par), nat.ptr(self.w_xx), nat.ptr(self.w_xh), nat.ptr(self.w_hh),
                 nat.ptr(self.H), nat.ptr(self.part), nat.ptr(self.tile_v), nat.ptr(self.enuc),
                 nat.ptr(self.enuc_pair), nat.ptr(b.pair_ref_device()) if enuc_pair else None, self.stream)

    # ------------------------------------------------------------------ fock
    def fock(self, P, F, active=None, eelec=None):
        b = self.b
        out = self.eelec if eelec is None else eelec
        nat.call("pyseqm_fock", b.L, nat.ptr(self.atom_par), nat.ptr(self.w_xx), nat.ptr(self.w_xh),
                 nat.ptr(self.w_hh), nat.ptr(P), nat.ptr(self.H), nat.ptr(F), nat.ptr(self.part),
                 nat.ptr(self.tile_v), nat.ptr(self.atom_v), nat.ptr(out), nat.ptr(active), self.stream)
        return out

    def initial_density(self):
        """Diagonal guess (scf_loop.py:2066-2081): heavy atoms tore/4 on s and p, H 1.0."""
        b = self.b
        P = torch.zeros(max(b.total, 1), dtype=torch.float64, device=self.device)
        didx, _ = b.diag_index()
        # diag_index is molecule-major in orbital order, the same order as the atoms
        k = np.where(b.Z > 1, 4, 1)
        vals = np.repeat(np.where(b.Z > 1, np.take(TORE, b.Z) / 4.0, 1.0), k)
        P[torch.as_tensor(didx, device=self.device)] = torch.as_tensor(vals, dtype=torch.float64,
                                                                        device=self.device)
        return P

    # ---------------------------------------------------------- density step
    def density_eigh(self, F, idx, nact, Pout, want_evecs=False):
        """FP64 density step by diagonalisation of the active Fock blocks -> Pout (diag.py:110-241).

        Blocks of up to EIGH_WG_MAX_N orbitals go to the one-workgroup-per-molecule solver
        (pyseqm_eigh_density: Householder with the block in registers, multisection, inverse
        iteration, cluster Gram-Schmidt, no padding) in one launch.  Larger blocks, and calls that
        want the eigenvectors themselves, take the batched
        library eigensolver on Gershgorin-padded stacks, again per size group (the reference pads
        every molecule to the batch's largest, which for an uneven batch multiplies the O(n^3) work)."""
        b = self.b
        if not want_evecs:
            idx_h = np.asarray(self._idx_host)
            wg = eigh_wg_ok(b.norb[idx_h], b.nocc[idx_h])
            if wg.any():
                # one launch, largest blocks first: the long solves start at once and the short
                # ones fill the CUs behind them (C2's density step: 5.2 ms against 8.4 ms for
                # launches per size group, each of which waits for its own largest block)
                ids = idx_h[wg]
                self._eigh_wg(F, ids[np.argsort(-b.norb[ids], kind="stable")], Pout)
            rest = idx_h[~wg]
            if len(rest) == 0:
                return None
            groups = size_groups(b.norb[rest])
            saved = self._idx_host
            for sel in groups:
                sub = torch.as_tensor(rest[sel].astype(np.int32), device=self.device)
                self._idx_host = rest[sel]
                self.density_eigh(F, sub, len(sel), Pout, want_evecs=True)
            self._idx_host = saved
            return None
        nmax = int(b.norb[self._idx_host].max())
        if nact >= EIGH_MIN_N_BATCH:
            nmax = max(nmax, EIGH_MIN_N)
        Fp = torch.empty(nact, nmax, nmax, dtype=torch.float64, device=self.device)
        nat.call("pyseqm_pad_for_eigh", b.L, nat.ptr(F), nat.ptr(idx), nact, nmax, nat.ptr(Fp), nat.ptr(self.red),
                 nat.red_blocks(nmax), self.stream)
        self.eigh_log.append((nact, nmax, "library"))
        e, V = torch.linalg.eigh(Fp)
        V = V.contiguous()
        nat.call("pyseqm_density_from_evecs", b.L, nat.ptr(V), nmax, nat.ptr(idx), nact, nat.ptr(Pout),
                 self.stream)
        if want_evecs:
            return e, V, nmax
        return None

    def _eigh_wg(self, F, ids, Pout):
        """One launch of the workgroup eigensolver on the molecules `ids` (host int array)."""
        b = self.b
        nmax = int(b.norb[ids].max())
        nocc_max = int(b.nocc[ids].max())
        ws = torch.empty(max(4 * len(ids) * nmax * nmax, 1), dtype=torch.float64, device=self.device)
        idx = torch.as_tensor(ids.astype(np.int32), device=self.device)
        nat.call("pyseqm_eigh_density", b.L, nat.ptr(idx), len(ids), nmax, nocc_max, nat.ptr(F), nat.ptr(Pout),
                 nat.ptr(ws), nat.ptr(self.eigh_fail), self.stream)
        self.eigh_log.append((len(ids), nmax, "workgroup"))

    def _sp2_workspace(self):
        ws = getattr(self, "_sp2ws", None)
        if ws is None:
            from .sp2 import SP2Workspace
            ws = self._sp2ws = SP2Workspace(self)
        return ws

    def density(self, F, idx, nact, Pout, fixed_iters=None):
        if self.density_kind == "eigh":
            self.density_eigh(F, idx, nact, Pout)
        else:
            from . import sp2
            sp2.density_sp2(self, F, idx, nact, Pout, fixed_iters=fixed_iters)

    # ------------------------------------------------------------------- SCF
    def _active(self, notconv_host):
        idx = np.nonzero(notconv_host)[0]
        self._idx_host = idx
        return torch.as_tensor(idx.astype(np.int32), device=self.device), len(idx)

    def scf(self, eps=1e-8, converger=(2,), density="eigh", sp2_eps=1e-4, P=None, max_iter=MAX_ITER, diis_tol=None):
        """scf_loop (scf_loop.py:2034-2394) with mixers [0] linear, [1] adaptive, [2] Pulay DIIS.
        `diis_tol` overrides the max|FP - PF| convergence test (default below)."""
        b = self.b
        dev = self.device
        self.density_kind = density
        self.sp2_eps = sp2_eps
        # max|FP - PF| test: 50 eps as the reference (scf_loop.py:52); the bf16x3 SP2 density has a
        # commutator floor of ~1e-3 (its energy error stays ~1e-8 eV/atom), so it is tested at 1e-2
        self.diis_tol = 50.0 * eps if density == "eigh" else max(50.0 * eps, SP2_DIIS_TOL)
        if diis_tol is not None:
            self.diis_tol = float(diis_tol)
        f64 = dict(dtype=torch.float64, device=dev)
        P = self.initial_density() if P is None else P.clone()
        self._metrics = torch.zeros(b.nmol, 4, **f64)
        F = torch.empty_like(P)
        Pnew = torch.zeros_like(P)
        Pold = torch.zeros_like(P)
        e_old = torch.zeros(b.nmol, **f64)
        e_new = torch.zeros(b.nmol, **f64)
        self.fock(P, F, eelec=e_old)
        notconv = torch.ones(b.nmol, dtype=torch.uint8, device=dev)
        nc_host = np.ones(b.nmol, dtype=bool)
        kind = converger[0]
        it = 0
        if kind == 2:
            res = self._pulay(P, F, Pnew, Pold, e_old, e_new, notconv, eps, max_iter)
        elif kind in (0, 1):
            res = self._simple(kind, converger, P, F, Pnew, Pold, e_old, e_new, notconv, eps, max_iter)
        else:
            raise ValueError("scf_converger must be [0, alpha], [1] or [2]")
        P, it, nc_host = res
        if density == "eigh":
            # the workgroup eigensolver counts eigenvectors it could not normalise (lost to their
            # cluster's span); one read per SCF
            lost = int(self.eigh_fail.item())
            if lost:
                self.eigh_fail.zero_()
                warnings.warn(f"FP64 eigensolver: {lost} eigenvector(s) lost to a cluster's span during this SCF; "
                              "the densities of the molecules involved may be inaccurate")
        out = SCFResult()
        out.P = P
        out.iterations = it
        out.notconverged = nc_host
        return out

    def _simple(self, kind, converger, P, F, Pnew, Pold, e_old, e_new, notconv, eps, max_iter):
        b = self.b
        nc_host = np.ones(b.nmol, dtype=bool)
        it = 0
        if kind == 1:
            dev = self.device
            old2 = torch.zeros(max(int(b.norb.sum()), 1), dtype=torch.float64, device=dev)  # diag of P two steps back
            orb_off = torch.as_tensor(np.concatenate([[0], np.cumsum(b.norb)[:-1]]).astype(np.int64), device=dev)
            fac = torch.zeros(b.nmol, dtype=torch.float64, device=dev)
            mask = torch.zeros(1, dtype=torch.int32, device=dev)
        krange = range(0, max_iter + 1) if kind == 0 else range(1, max_iter + 1)
        for k in krange:
            idx, nact = self._active(nc_host)
            if nact == 0:
                break
            it += 1
            nmax = int(b.norb[self._idx_host].max())
            self.density(F, idx, nact, Pnew)
            if kind == 0:
                nat.call("pyseqm_mix", b.L, nat.ptr(idx), nact, nmax, float(converger[1]), nat.ptr(P),
                         nat.ptr(Pnew), nat.ptr(Pold), self.stream)
            else:
                nat.call("pyseqm_adaptive_mix", b.L, nat.ptr(idx), nact, nmax, int(k), nat.ptr(P), nat.ptr(Pnew),
                         nat.ptr(Pold), nat.ptr(old2), nat.ptr(orb_off), nat.ptr(fac), nat.ptr(mask), self.stream)
            act = self._act_mask(nc_host)
            self.fock(P, F, active=act, eelec=e_new)
            nat.call("pyseqm_scf_check", b.L, nat.ptr(idx), nact, nat.ptr(P), nat.ptr(Pold), nat.ptr(e_new),
                     nat.ptr(e_old), None, float(eps), 0.0, nat.ptr(notconv), None, nat.ptr(self.red),
                     nat.red_blocks(nmax), self.stream)
            nc_host = notconv.cpu().numpy().astype(bool)
        return P, it, nc_host

    def _act_mask(self, nc_host):
        return torch.as_tensor(nc_host.astype(np.uint8), device=self.device)

    def _pulay(self, P, F, Pnew, Pold, e_old, e_new, notconv, eps, max_iter):
        """scf_forward2 (scf_loop.py:639-1132): Pulay DIIS on FP - PF with a 10-deep history,
        batch-global history counter and reset on ill-conditioning (cond > 1e7)."""
        b = self.b
        dev = self.device
        total = max(b.total, 1)
        f64 = dict(dtype=torch.float64, device=dev)
        Fring = torch.zeros(NFOCK, total, **f64)
        Cring = torch.zeros(NFOCK, total, **f64)
        X = torch.empty(total, **f64)
        diis_err = torch.full((b.nmol,), torch.finfo(torch.float64).max, **f64)
        emat0 = torch.tril(torch.eye(NFOCK + 1, **f64) - 1.0)
        # clone, not contiguous(): for one molecule the expanded view already counts as contiguous and
        # would alias emat0, so the reset below would copy the updated matrix onto itself
        EMAT = emat0.expand(b.nmol, -1, -1).clone()
        coef = torch.zeros(b.nmol, NFOCK, **f64)
        reset_flags = torch.zeros(b.nmol, dtype=torch.uint8, device=dev)
        nc_host = np.ones(b.nmol, dtype=bool)
        counter, cF = -1, 0
        it = 0
        for k in range(max_iter + 1):
            idx, nact = self._active(nc_host)
            if nact == 0:
                break
            it += 1
            nmax = int(b.norb[self._idx_host].max())
            cF = min(cF + 1, NFOCK)
            counter = (counter + 1) % NFOCK
            Fslot = Fring[counter]
            nat.call("pyseqm_copy_active", b.L, nat.ptr(idx), nact, nmax, nat.ptr(F), nat.ptr(Fslot), self.stream)
            nb = nat.red_blocks(nmax)
            if self.density_kind == "sp2" and nmax > MIXED_DIIS_MIN_N:
                # bf16x3 commutator on the upper triangle (the ring is read there only), from FP32
                # copies of F and P in the SP2 X buffers, which are free until the density step
                ws = self._sp2_workspace()
                nat.call("pyseqm_diis_commutator_bf16x3", b.L, nat.ptr(idx), nact, nmax, nat.ptr(F), nat.ptr(P),
                         nat.ptr(ws.xoff), nat.ptr(ws.xld), nat.ptr(ws.Xa), nat.ptr(ws.Xb), nat.ptr(Cring[counter]),
                         nat.ptr(diis_err), nat.ptr(self.red), nat.MAX_RED_BLOCKS * 16, self.stream)
            else:
                nat.call("pyseqm_diis_residual", b.L, nat.ptr(idx), nact, nmax, nat.ptr(F), nat.ptr(P), nat.ptr(X),
                         nat.ptr(Cring[counter]), nat.ptr(diis_err), nat.ptr(self.red), nb, 0, self.stream)
            nat.call("pyseqm_diis_dots", b.L, nat.ptr(idx), nact, nat.ptr(Cring), total, counter, cF, NFOCK,
                     nat.ptr(EMAT), nat.ptr(self.red), nb, self.stream)
            reset_t = None
            if cF >= 2:
                # the reference's eigh reads only the lower triangle (stale upper entries included):
                # the kernel mirrors it, normalises by B[slot][slot] and solves by FP64 Jacobi
                reset_flags.zero_()
                nat.call("pyseqm_diis_solve", b.L, nat.ptr(idx), nact, NFOCK, cF, counter, nat.ptr(EMAT),
                         nat.ptr(coef), nat.ptr(reset_flags), self.stream)
                reset_t = reset_flags.any()  # read with the convergence flags (one host sync)
                if self.diis_log is not None:
                    self.diis_log.append((cF, counter, EMAT.cpu().numpy().copy(), coef.cpu().numpy().copy(),
                                          reset_flags.cpu().numpy().copy(), diis_err.cpu().numpy().copy()))
                nat.call("pyseqm_diis_combine", b.L, nat.ptr(idx), nact, nmax, nat.ptr(Fring), total, cF,
                         nat.ptr(coef), NFOCK, nat.ptr(F), self.stream)
            self.density(F, idx, nact, Pnew)
            alpha = 0.5 if cF < 2 else 0.0
            nat.call("pyseqm_mix", b.L, nat.ptr(idx), nact, nmax, alpha, nat.ptr(P), nat.ptr(Pnew),
                     nat.ptr(Pold), self.stream)
            act = self._act_mask(nc_host)
            self.fock(P, F, active=act, eelec=e_new)
            nat.call("pyseqm_scf_check", b.L, nat.ptr(idx), nact, nat.ptr(P), nat.ptr(Pold), nat.ptr(e_new),
                     nat.ptr(e_old), nat.ptr(diis_err), float(eps), self.diis_tol, nat.ptr(notconv),
                     nat.ptr(self._metrics), nat.ptr(self.red), nb, self.stream)
            nc_host = notconv.cpu().numpy().astype(bool)
            reset = bool(reset_t.item()) if reset_t is not None else False
            if self.trace is not None:
                self.trace.append(self._metrics.cpu().numpy().copy())
            if reset:
                counter, cF = -1, 0
                EMAT.copy_(emat0.expand(b.nmol, -1, -1))
        return P, it, nc_host

    # -------------------------------------------------------------- results
    def energies(self, P):
        """Final F = fock(P), Eelec, Enuc, Etot, Eiso, Hf per molecule (basics.py:862-892, 1220-1239)."""
        F = torch.empty_like(P)
        eelec = torch.zeros(self.b.nmol, dtype=torch.float64, device=self.device)
        self.fock(P, F, eelec=eelec)
        etot = eelec + self.enuc
        hf = etot - self.eiso + self.eheat
        return dict(F=F, Eelec=eelec, Enuc=self.enuc.clone(), Etot=etot, Eiso=self.eiso.clone(), Hf=hf)

    def forces(self, P, h=1.0e-3):
        """Forces (natoms, 3) in eV/Angstrom at fixed density P (Hellmann-Feynman, the reference's
        autograd force with scf_backward = 0, basics.py:1315-1336)."""
        b = self.b
        out = torch.zeros(max(b.natoms, 1), 3, dtype=torch.float64, device=self.device)
        nat.call("pyseqm_pair_forces", b.L, nat.METHODS[self.method], nat.ptr(self.coords), nat.ptr(self.atom_par),
                 nat.ptr(P), float(h), nat.ptr(self.part), nat.ptr(out), self.stream)
        return out[:b.natoms]

    def param_grad(self, P, hrel=1.0e-4, atom_par=None, coords=None):
        """Derivatives of the energy at fixed density P with respect to every atom's 24 input
        parameters (COLUMNS order): (natoms, 2, 24), [:, 0] for Eelec and [:, 1] for the core
        repulsion Enuc.  At a converged SCF density this is the exact gradient of Etot (the energy is
        stationary in P), the one the reference's autograd returns for learned parameters with
        scf_backward = 0 (basics.py:781-789, scf_loop.py:1996-2031).  Pair terms by 4th-order
        central differences of each pair's energy (step hrel * max(|value|, 0.1)), one-center terms
        in closed form.  atom_par / coords default to the engine's current ones."""
        b = self.b
        out = torch.zeros(max(b.natoms, 1), 2, 24, dtype=torch.float64, device=self.device)
        rec = torch.empty(max(int(nat.lib().pyseqm_param_grad_workspace(nat.ctypes.byref(b.L))), 1),
                          dtype=torch.float64, device=self.device)
        ap = self.atom_par if atom_par is None else atom_par.contiguous()
        xyz = self.coords if coords is None else coords.contiguous()
        nat.call("pyseqm_param_grad", b.L, nat.METHODS[self.method], nat.ptr(xyz), nat.ptr(ap),
                 nat.ptr(P), float(hrel), nat.ptr(rec), nat.ptr(self.part), nat.ptr(out), self.stream)
        return out[:b.natoms]

    def eiso_grad(self):
        """d(isolated-atom energy)/d(input parameter) per atom, (natoms, 24): the coefficients of
        U_ss, U_pp, g_ss, g_sp, g_pp, g_p2, h_sp in Eiso (energy.py:8-23)."""
        b = self.b
        Z = torch.as_tensor(b.Z, device=self.device).long()
        tab = lambda t: torch.as_tensor(t, dtype=torch.float64, device=self.device)[Z]
        g = torch.zeros(b.natoms, 24, dtype=torch.float64, device=self.device)
        for name, coef in (("U_ss", USSC), ("U_pp", UPPC), ("g_ss", GSSC), ("g_sp", GSPC), ("g_pp", GPPC),
                           ("g_p2", GP2C), ("h_sp", HSPC)):
            g[:, COLUMNS.index(name)] = tab(coef)
        return g

    def eigenpairs(self, F, orb_off, evals, Q, eig_ws, fail):
        """Every eigenpair of every block of F: Q in the packed layout (column j = eigenvector j,
        ascending eigenvalues) and evals[orb_off[m] + j].  Blocks whose full decomposition fits the
        workgroup eigensolver's LDS (at most 139 orbitals) take pyseqm_eigh_vectors; larger ones the
        batched library eigensolver on Gershgorin-padded stacks per size group, scattered into the
        same layout.  eig_ws holds 4 nmol n^2 doubles for the n of the workgroup blocks; fail counts
        vectors the workgroup solver lost (read by the caller)."""
        b = self.b
        norb = np.asarray(b.norb)
        wg = eigh_wg_ok(norb, norb)
        ids = np.nonzero(wg)[0]
        if len(ids):
            idx = torch.as_tensor(ids.astype(np.int32), device=self.device)
            nat.call("pyseqm_eigh_vectors", b.L, nat.ptr(idx), len(ids), int(norb[ids].max()), nat.ptr(F),
                     nat.ptr(orb_off), nat.ptr(evals), nat.ptr(Q), nat.ptr(eig_ws), nat.ptr(fail), self.stream)
        rest = np.nonzero(~wg)[0]
        if len(rest) == 0:
            return
        off_h = np.concatenate([[0], np.cumsum(norb)[:-1]]).astype(np.int64)
        for sel in size_groups(norb[rest]):
            g = rest[sel]
            nmax = int(norb[g].max())
            idx = torch.as_tensor(g.astype(np.int32), device=self.device)
            Fp = torch.empty(len(g), nmax, nmax, dtype=torch.float64, device=self.device)
            nat.call("pyseqm_pad_for_eigh", b.L, nat.ptr(F), nat.ptr(idx), len(g), nmax, nat.ptr(Fp),
                     nat.ptr(self.red), nat.red_blocks(nmax), self.stream)
            e0, V = torch.linalg.eigh(Fp)  # padded eigenvalues lie above each block's spectrum
            for k, m in enumerate(g):
                n = int(norb[m])
                o = int(b.mol_mat_off[m])
                Q[o:o + n * n].view(n, n).copy_(V[k, :n, :n])
                evals[int(off_h[m]):int(off_h[m]) + n].copy_(e0[k, :n])

    def eig(self, F):
        """Orbital energies in the reference layout e (nmol, 4*molsize), zero padded, and the packed
        eigenvectors (nmol, nmax, nmax), sym_eig_trunc(eig_only=True) (diag.py:110-241)."""
        b = self.b
        self._idx_host = np.arange(b.nmol)
        idx = torch.arange(b.nmol, dtype=torch.int32, device=self.device)
        nmax = b.nmax
        Fp = torch.empty(b.nmol, nmax, nmax, dtype=torch.float64, device=self.device)
        nat.call("pyseqm_pad_for_eigh", b.L, nat.ptr(F), nat.ptr(idx), b.nmol, nmax, nat.ptr(Fp),
                 nat.ptr(self.red), nat.red_blocks(nmax), self.stream)
        e0, V = torch.linalg.eigh(Fp)
        e = torch.zeros(b.nmol, 4 * b.molsize, dtype=torch.float64, device=self.device)
        mask = torch.arange(nmax, device=self.device)[None, :] < torch.as_tensor(b.norb, device=self.device)[:, None]
        e[:, :nmax] = torch.where(mask, e0, torch.zeros_like(e0))
        return e, V

    # ------------------------------------------------- reference-order views
    def w_reference(self):
        """Two-center integrals as the reference's dense (npairs, 10, 10) array (hcore.py:9-179)."""
        b = self.b
        dev = self.device
        w = torch.zeros(max(b.npairs, 1), 10, 10, dtype=torch.float64, device=dev)
        for kind, buf in enumerate((self.w_xx, self.w_xh, self.w_hh)):
            ref, idx = b.w_gather_index(kind, dev)
            if ref is None:
                continue
            rows = buf[idx]
            if kind == 0:
                w[ref] = rows.view(-1, 10, 10)
            elif kind == 1:
                w[ref, :, 0] = rows
            else:
                w[ref, 0, 0] = rows[:, 0]
        return w[:b.npairs]
