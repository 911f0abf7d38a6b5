"""The C ABI registered as PyTorch custom operators (namespace ``pyseqm``).

Each operator wraps one ``extern "C"`` entry point of libpyseqm_hip.so: it takes device tensors
(outputs are mutated in place, as the C functions write into caller buffers), runs on the
current HIP stream and returns nothing.  The packed-batch descriptor (``pyseqm_layout_t``) is
passed as an integer handle from :func:`layout_handle`.  These are the same kernels the engine
calls; the engine's inner SCF loop calls the C ABI directly to avoid the dispatcher's per-call
cost on its many small launches.

    h = layout_handle(batch)
    torch.ops.pyseqm.fock_build(h, atom_par, w_xx, w_xh, w_hh, P, H, F, part, tile_v, atom_v, eelec)
"""
import torch

from . import _native as nat

_LAYOUTS = {}


def layout_handle(batch):
    """Integer handle of a PackedBatch / MatrixLayout for the operators below (kept alive here)."""
    key = id(batch)
    _LAYOUTS[key] = batch
    return key


def release_handle(key):
    _LAYOUTS.pop(key, None)


def _L(h):
    try:
        return _LAYOUTS[h].L
    except KeyError:
        raise ValueError(f"unknown pyseqm layout handle {h}") from None


def _s():
    return nat.stream_ptr()


def _p(t):
    return nat.ptr(t)


def _check_nmax(layout, idx, nmax, V=None):
    """Every listed block must fit the padded nmax x nmax stack (V: nact blocks of nmax^2)."""
    b = _LAYOUTS[layout] if layout in _LAYOUTS else None
    if b is None:
        raise ValueError(f"unknown pyseqm layout handle {layout}")
    ids = idx.detach().cpu().numpy()
    if len(ids) and int(b.norb[ids].max()) > nmax:
        raise ValueError(f"nmax = {nmax} is below the largest listed block ({int(b.norb[ids].max())} orbitals)")
    if V is not None and V.numel() < len(ids) * nmax * nmax:
        raise ValueError(f"V holds {V.numel()} values, fewer than {len(ids)} blocks of {nmax} x {nmax}")


@torch.library.custom_op("pyseqm::atom_setup", mutates_args=("atom_par",))
def atom_setup(layout: int, atom_par: torch.Tensor) -> None:
    """Derived multipole parameters per atom (two_elec_two_center_int.py:117-147, cal_par.py)."""
    nat.call("pyseqm_atom_setup", _L(layout), _p(atom_par), _s())


@torch.library.custom_op("pyseqm::pair_integrals", mutates_args=("w_xx", "w_xh", "w_hh", "H", "part", "tile_v",
                                                                  "enuc"))
def pair_integrals(layout: int, method: int, coords: torch.Tensor, atom_par: torch.Tensor, w_xx: torch.Tensor,
                   w_xh: torch.Tensor, w_hh: torch.Tensor, H: torch.Tensor, part: torch.Tensor, tile_v: torch.Tensor,
                   enuc: torch.Tensor) -> None:
    """Two-center integrals, core Hamiltonian and core-core repulsion (hcore.py:9-179)."""
    nat.call("pyseqm_pair_integrals", _L(layout), method, _p(coords), _p(atom_par), _p(w_xx), _p(w_xh), _p(w_hh),
             _p(H), _p(part), _p(tile_v), _p(enuc), None, None, _s())


@torch.library.custom_op("pyseqm::fock_build", mutates_args=("F", "part", "tile_v", "atom_v", "eelec"))
def fock_build(layout: int, atom_par: torch.Tensor, w_xx: torch.Tensor, w_xh: torch.Tensor, w_hh: torch.Tensor,
               P: torch.Tensor, H: torch.Tensor, F: torch.Tensor, part: torch.Tensor, tile_v: torch.Tensor,
               atom_v: torch.Tensor, eelec: torch.Tensor) -> None:
    """F = H + G(P) and Eelec (fock.py:132-347, energy.py:26-49)."""
    nat.call("pyseqm_fock", _L(layout), _p(atom_par), _p(w_xx), _p(w_xh), _p(w_hh), _p(P), _p(H), _p(F), _p(part),
             _p(tile_v), _p(atom_v), _p(eelec), None, _s())


@torch.library.custom_op("pyseqm::pad_for_eigh", mutates_args=("out", "red"))
def pad_for_eigh(layout: int, F: torch.Tensor, idx: torch.Tensor, nmax: int, out: torch.Tensor,
                 red: torch.Tensor) -> None:
    """Gershgorin-shifted padded stack of the listed Fock blocks (diag.py:130-160)."""
    _check_nmax(layout, idx, nmax, out)
    nat.call("pyseqm_pad_for_eigh", _L(layout), _p(F), _p(idx), int(idx.numel()), nmax, _p(out), _p(red),
             nat.red_blocks(nmax), _s())


@torch.library.custom_op("pyseqm::density_from_evecs", mutates_args=("P",))
def density_from_evecs(layout: int, V: torch.Tensor, nmax: int, idx: torch.Tensor, P: torch.Tensor) -> None:
    """P = 2 C_occ C_occ^T for the listed molecules (diag.py:230-232)."""
    _check_nmax(layout, idx, nmax, V)
    nat.call("pyseqm_density_from_evecs", _L(layout), _p(V), nmax, _p(idx), int(idx.numel()), _p(P), _s())


@torch.library.custom_op("pyseqm::pair_forces", mutates_args=("part", "forces"))
def pair_forces(layout: int, method: int, coords: torch.Tensor, atom_par: torch.Tensor, P: torch.Tensor, h: float,
                part: torch.Tensor, forces: torch.Tensor) -> None:
    """Hellmann-Feynman forces at fixed P (basics.py:1315-1336)."""
    nat.call("pyseqm_pair_forces", _L(layout), method, _p(coords), _p(atom_par), _p(P), float(h), _p(part),
             _p(forces), _s())


@torch.library.custom_op("pyseqm::pair_forces_xl", mutates_args=("part", "forces"))
def pair_forces_xl(layout: int, method: int, coords: torch.Tensor, atom_par: torch.Tensor, D: torch.Tensor,
                   P: torch.Tensor, h: float, part: torch.Tensor, forces: torch.Tensor) -> None:
    """XL-BOMD forces at fixed D, P (xlbomd.py:536-551)."""
    nat.call("pyseqm_pair_forces_xl", _L(layout), method, _p(coords), _p(atom_par), _p(D), _p(P), float(h),
             _p(part), _p(forces), _s())


@torch.library.custom_op("pyseqm::xl_energy", mutates_args=("eelec", "red"))
def xl_energy(layout: int, D: torch.Tensor, P: torch.Tensor, F: torch.Tensor, H: torch.Tensor, eelec: torch.Tensor,
              red: torch.Tensor) -> None:
    """sum D F - 1/2 sum (F - H) P per molecule (energy.py:76-88)."""
    nat.call("pyseqm_xl_energy", _L(layout), _p(D), _p(P), _p(F), _p(H), _p(eelec), _p(red),
             nat.red_blocks(_LAYOUTS[layout].nmax), _s())


@torch.library.custom_op("pyseqm::xl_propagate", mutates_args=("P", "Pt"))
def xl_propagate(D: torch.Tensor, P: torch.Tensor, Pt: torch.Tensor, coef: list[float], coef_D: float, c: float,
                 slot: int) -> None:
    """P <- coef_D (c D + (1-c) P) + sum_j coef[j] Pt[j];  Pt[slot] = P (MolecularDynamics.py:1418-1461)."""
    import ctypes
    arr = (ctypes.c_double * len(coef))(*coef)
    nat.call_raw("pyseqm_xl_propagate", int(P.numel()), _p(D), _p(P), _p(Pt), int(Pt.shape[1]), len(coef), arr,
                 float(coef_D), float(c), int(slot), _s())


@torch.library.custom_op("pyseqm::core_repulsion_pairs", mutates_args=("out",))
def core_repulsion_pairs(method: int, idxi: torch.Tensor, idxj: torch.Tensor, rij: torch.Tensor, gam: torch.Tensor,
                         atom_par: torch.Tensor, out: torch.Tensor) -> None:
    """Core-core repulsion of an explicit pair list (energy.py:91-134)."""
    nat.call_raw("pyseqm_core_repulsion_pairs", int(rij.numel()), method, _p(idxi), _p(idxj), _p(rij), _p(gam),
                 _p(atom_par), _p(out), _s())


@torch.library.custom_op("pyseqm::density_eigh_fp64", mutates_args=("P", "red"))
def density_eigh_fp64(layout: int, F: torch.Tensor, idx: torch.Tensor, nmax: int, P: torch.Tensor,
                      red: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """sym_eig_trunc (diag.py:110-241) on the listed blocks: Gershgorin-padded nmax x nmax stack, batched
    FP64 eigensolve, P = 2 C_occ C_occ^T into the packed density.  Returns the padded eigenvalues
    (nact, nmax) and eigenvectors (nact, nmax, nmax, columns)."""
    nact = int(idx.numel())
    _check_nmax(layout, idx, nmax)
    Fp = torch.empty(nact, nmax, nmax, dtype=torch.float64, device=F.device)
    nat.call("pyseqm_pad_for_eigh", _L(layout), _p(F), _p(idx), nact, nmax, _p(Fp), _p(red), nat.red_blocks(nmax),
             _s())
    e, V = torch.linalg.eigh(Fp)
    V = V.contiguous()
    nat.call("pyseqm_density_from_evecs", _L(layout), _p(V), nmax, _p(idx), nact, _p(P), _s())
    return e, V


@torch.library.custom_op("pyseqm::density_sp2", mutates_args=("Xa", "Xb", "state", "part", "Y2", "red", "ndone", "P"))
def density_sp2(layout: int, F: torch.Tensor, idx: torch.Tensor, nmax: int, xoff: torch.Tensor, xld: torch.Tensor,
                Xa: torch.Tensor, Xb: torch.Tensor, state: torch.Tensor, part: torch.Tensor, Y2: torch.Tensor,
                red: torch.Tensor, ndone: torch.Tensor, P: torch.Tensor, eps: float, iters: int) -> None:
    """SP2 (SP2.py:9-85) in mixed precision on the listed blocks, then the McWeeny step (FP64 accuracy) into the packed
    density P.  `iters` iterations are launched (molecules that stop earlier make the rest no-ops), so the
    call makes no host synchronisation.  Buffers as pyseqm_amd.sp2.SP2Workspace lays them out: X blocks
    (FP32, Xa / Xb) at xoff with leading dimension xld, state (nmol, 8), part (nmol, part_stride), Y2 (FP64,
    X layout), ndone (1 int)."""
    L = _L(layout)
    nact = int(idx.numel())
    nat.call("pyseqm_sp2_init", L, _p(F), _p(idx), nact, _p(xoff), _p(xld), _p(Xa), _p(state), _p(red),
             nat.red_blocks(nmax), _s())
    ndone.zero_()
    nat.call("pyseqm_sp2_iterate_many", L, _p(idx), nact, nmax, _p(xoff), _p(xld), _p(Xa), _p(Xb), _p(state),
             _p(part), int(part.shape[-1]), float(eps), int(iters), _p(ndone), 0, int(iters), _s())
    nat.call("pyseqm_sp2_mcweeny", L, _p(idx), nact, nmax, _p(xoff), _p(xld), _p(state), _p(Xa), _p(Xb), _p(Y2),
             _p(P), _s())


@torch.library.custom_op("pyseqm::mix", mutates_args=("P", "Pold"))
def mix(layout: int, idx: torch.Tensor, nmax: int, alpha: float, P: torch.Tensor, Pnew: torch.Tensor,
        Pold: torch.Tensor) -> None:
    """Pold = P; P = alpha P + (1 - alpha) Pnew on the listed blocks (scf_loop.py:266-267, 1052-1056)."""
    nat.call("pyseqm_mix", _L(layout), _p(idx), int(idx.numel()), nmax, float(alpha), _p(P), _p(Pnew), _p(Pold),
             _s())


@torch.library.custom_op("pyseqm::scf_check", mutates_args=("e_old", "notconv", "metrics", "red"))
def scf_check(layout: int, idx: torch.Tensor, nmax: int, P: torch.Tensor, Pold: torch.Tensor, e_new: torch.Tensor,
              e_old: torch.Tensor, diis_err: torch.Tensor, eps: float, diis_tol: float, notconv: torch.Tensor,
              metrics: torch.Tensor, red: torch.Tensor) -> None:
    """The 4-way convergence test get_error (scf_loop.py:106-147) per listed molecule."""
    nat.call("pyseqm_scf_check", _L(layout), _p(idx), int(idx.numel()), _p(P), _p(Pold), _p(e_new), _p(e_old),
             _p(diis_err), float(eps), float(diis_tol), _p(notconv), _p(metrics), _p(red), nat.red_blocks(nmax), _s())


@torch.library.custom_op("pyseqm::diis_step", mutates_args=("Fring", "Cring", "X", "diis_err", "Bmat", "coef",
                                                             "reset", "F", "red"))
def diis_step(layout: int, idx: torch.Tensor, nmax: int, F: torch.Tensor, P: torch.Tensor, Fring: torch.Tensor,
              Cring: torch.Tensor, X: torch.Tensor, diis_err: torch.Tensor, Bmat: torch.Tensor, coef: torch.Tensor,
              reset: torch.Tensor, red: torch.Tensor, slot: int, count: int, mixed: bool) -> None:
    """One Pulay/DIIS update (scf_loop.py:990-1035) on the listed blocks: F into ring slot `slot`, the
    residual FP - PF into the error ring with its max-norm, the new B-matrix row against the `count`
    stored entries and, when count >= 2, the coefficients (FP64 Jacobi solve with the reference's
    pseudo-inverse and reset flag) and the extrapolated F.  Fring / Cring are (nfock, total) rings."""
    L = _L(layout)
    nact = int(idx.numel())
    nfock = int(Fring.shape[0])
    total = int(Fring.shape[1])
    nb = nat.red_blocks(nmax)
    nat.call("pyseqm_copy_active", L, _p(idx), nact, nmax, _p(F), _p(Fring[slot]), _s())
    nat.call("pyseqm_diis_residual", L, _p(idx), nact, nmax, _p(F), _p(P), _p(X), _p(Cring[slot]), _p(diis_err),
             _p(red), nb, int(bool(mixed)), _s())
    nat.call("pyseqm_diis_dots", L, _p(idx), nact, _p(Cring), total, slot, count, nfock, _p(Bmat), _p(red), nb, _s())
    if count >= 2:
        reset.zero_()
        nat.call("pyseqm_diis_solve", L, _p(idx), nact, nfock, count, slot, _p(Bmat), _p(coef), _p(reset), _s())
        nat.call("pyseqm_diis_combine", L, _p(idx), nact, nmax, _p(Fring), total, count, _p(coef), nfock, _p(F), _s())


OPS = ("atom_setup", "pair_integrals", "fock_build", "pad_for_eigh", "density_from_evecs", "pair_forces",
       "pair_forces_xl", "xl_energy", "xl_propagate", "core_repulsion_pairs", "density_eigh_fp64", "density_sp2",
       "mix", "scf_check", "diis_step")
