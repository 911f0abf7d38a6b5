"""scf_loop: the reference's SCF entry point (seqm/seqm_functions/scf_loop.py:2034-2394) on the
packed device engine.  Returns the reference's 12-tuple in the reference's padded layouts."""
import warnings

import torch

from ..engine import MAX_ITER

RAISE_ERROR_IF_SCF_FORWARD_FAILS = False


def _converger(scf_converger):
    kind = int(scf_converger[0])
    if kind == 0:
        return (0, float(scf_converger[1]) if len(scf_converger) > 1 else 0.5)
    if kind in (1, 2):
        return (kind,)
    raise NotImplementedError("scf_converger [3] (KSA) is outside the closed-shell sp SCF path")


def orbital_charges(molecule, v):
    """Per-orbital charge on each atom from the eigenvectors (scf_loop.py:2376-2387)."""
    nmol = molecule.nmol
    m = molecule.molsize
    charge = torch.zeros(nmol, 4 * m, m, dtype=v.dtype, device=v.device)
    v2 = v * v
    for i in range(nmol):
        nh, nH = int(molecule.nHeavy[i]), int(molecule.nHydro[i])
        n = 4 * nh + nH
        charge[i, :n, :nh] = v2[i][:n, :4 * nh].reshape(n, 4, nh).sum(dim=1)
        charge[i, :n, nh:nh + nH] = v2[i][:n, 4 * nh:4 * nh + nH]
    return charge


def _check_pair_cutoff(molecule):
    """The device engine couples every atom pair of a molecule; a pair_outer_cutoff that would drop
    pairs (the reference's Parser keeps only r < cutoff, basics.py:313-328) is refused."""
    cut = float(molecule.seqm_parameters.get("pair_outer_cutoff", 1.0e10))
    if cut >= 1.0e9:
        return
    x = molecule.coordinates.detach()
    real = molecule.species > 0
    d = torch.cdist(x, x)
    d = torch.where(real[:, :, None] & real[:, None, :], d, torch.zeros_like(d))
    if bool((d >= cut).any()):
        raise NotImplementedError(f"pair_outer_cutoff = {cut} A drops atom pairs; the device engine uses every "
                                  "pair of a molecule (leave the cutoff unset or above the largest distance)")


def scf_loop(molecule, eps=1.0e-4, P=None, sp2=[False], scf_converger=[1], eig=False, scf_backward=0,
             scf_backward_eps=1.0e-2):
    if scf_backward not in (0, 1, 2):
        raise ValueError(f"scf_backward must be 0, 1 or 2 (got {scf_backward!r})")
    _check_pair_cutoff(molecule)
    eng = molecule.engine()
    eng.set_coordinates(molecule.coordinates.detach())
    b = eng.b
    P0 = b.from_padded(P.detach().to(torch.float64)) if torch.is_tensor(P) else None
    eps = float(eps)
    density = "sp2" if sp2[0] else "eigh"
    sp2_eps = float(sp2[1]) if (sp2[0] and len(sp2) > 1) else 1.0e-2
    r = eng.scf(eps=eps, converger=_converger(scf_converger), density=density, sp2_eps=sp2_eps, P=P0,
                max_iter=MAX_ITER)
    nc = torch.as_tensor(r.notconverged, device=molecule.coordinates.device)
    if nc.any():
        nnot = int(nc.sum())
        warnings.warn("SCF for %d/%d molecules doesn't converge after %d iterations" % (nnot, b.nmol, MAX_ITER))
        if RAISE_ERROR_IF_SCF_FORWARD_FAILS:
            raise ValueError("SCF for some the molecules in the batch doesn't converge")
    eng._last_P = r.P
    en = eng.energies(r.P)
    F = b.to_padded(en["F"])
    Ppad = b.to_padded(r.P)
    Hcore = torch.triu(b.to_padded(eng.H))  # the reference fills only the upper triangle
    w = eng.w_reference()
    rho0 = eng.atom_par[:, 26]
    rho0xi = rho0[molecule.idxi]
    rho0xj = rho0[molecule.idxj]
    e = V = charge = None
    if eig:
        e, V = eng.eig(en["F"])
        charge = orbital_charges(molecule, V)
    if scf_backward >= 1:
        # the density (and the orbital energies) carry the gradient of the SCF fixed point with
        # respect to the learned parameters (scf_loop.py:1779-1993); 2 (the reference's unrolled
        # iterations) gives the same gradient at convergence and takes the same implicit route here
        from ..basics import _trained_parameters
        trained = _trained_parameters(molecule.learned_parameters)
        if trained:
            from ..implicit import SCFImplicit
            out = SCFImplicit.apply(eng, r.P, en["F"], float(scf_backward_eps), r.notconverged,
                                    [n for n, _ in trained], Ppad, e, *[t for _, t in trained])
            Ppad, e = out if eig else (out, None)
    return F, e, Ppad, Hcore, w, charge, rho0xi, rho0xj, None, None, nc, V
