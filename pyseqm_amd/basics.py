"""Parser / Pack_Parameters / Hamiltonian / Energy / Force of the drop-in API (seqm/basics.py).

The classes keep the reference's constructor arguments, call signatures and return tuples;
the numerical work runs in the packed device engine (``SCFEngine``): pair integrals, Fock
builds, density steps, energies and forces are HIP kernels.
"""
import numpy as np
import torch

from .seqm_functions.constants import ev
from .seqm_functions.parameters import COLUMNS, METHODS, param_table

parameterlist = {
    "MNDO": COLUMNS[:12],
    "AM1": COLUMNS[:12] + [f"Gaussian{i}_{x}" for x in "KLM" for i in range(1, 5)],
    "PM3": COLUMNS[:12] + [f"Gaussian{i}_{x}" for x in "KLM" for i in range(1, 3)],
}


class Parser(torch.nn.Module):
    """Pair list and per-atom indices (seqm/basics.py:199-403): returns the reference's tuple
    (nmol, molsize, nSuperHeavy, nHeavy, nHydro, nocc, Z, maskd, atom_molid, mask, [mask_l,]
    pair_molid, ni, nj, idxi, idxj, xij, rij)."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.outercutoff = seqm_parameters.get("pair_outer_cutoff", 1e10)
        if seqm_parameters.get("elements") is None:
            raise RuntimeError("Please instantiate a Molecule object before instantiating an object of "
                               "Electronic_structure. Alternately, provide a list called of unique elements in "
                               "the input molecules under the 'elements' key in seqm_parameters")
        self.elements = seqm_parameters["elements"]

    def forward(self, molecule, themethod, return_mask_l=False, do_large_tensors=True, *args, **kwargs):
        dev = molecule.coordinates.device
        species = molecule.species
        nmol, molsize = species.shape
        nonblank = species > 0
        flat = torch.nonzero(nonblank.reshape(-1), as_tuple=False).squeeze(1)
        Z = species.reshape(-1)[flat]
        nHeavy = (species > 1).sum(dim=1)
        nHydro = (species == 1).sum(dim=1)
        nSuperHeavy = torch.zeros_like(nHeavy)
        n_el = molecule.const.tore[species].sum(dim=1).to(torch.int64)
        if torch.is_tensor(molecule.tot_charge):
            n_el = n_el - molecule.tot_charge.reshape(-1).to(torch.int64)
        if ((n_el % 2) == 1).any():
            raise ValueError("RHF setting requires closed shell systems (even number of electrons)")
        nocc = n_el // 2
        maskd = (torch.arange(nmol, device=dev)[:, None] * molsize ** 2
                 + torch.arange(molsize, device=dev)[None, :] * (molsize + 1)).reshape(-1)[flat]
        atom_molid = flat // molsize
        iu, ju = torch.triu_indices(molsize, molsize, 1, device=dev)
        pm = torch.arange(nmol, device=dev).repeat_interleave(len(iu))
        pi, pj = iu.repeat(nmol), ju.repeat(nmol)
        xyz = molecule.coordinates.detach()
        d = xyz[pm, pj] - xyz[pm, pi]
        dist = torch.sqrt((d * d).sum(1))
        keep = nonblank[pm, pi] & nonblank[pm, pj] & (dist < self.outercutoff)
        pm, pi, pj, d, dist = pm[keep], pi[keep], pj[keep], d[keep], dist[keep]
        inv = torch.full((nmol * molsize,), -1, dtype=torch.int64, device=dev)
        inv[flat] = torch.arange(len(flat), device=dev)
        idxi = inv[pm * molsize + pi]
        idxj = inv[pm * molsize + pj]
        ni, nj = Z[idxi], Z[idxj]
        xij = d / dist.unsqueeze(1)
        rij = dist * molecule.const.length_conversion_factor
        mask = pm * molsize ** 2 + pi * molsize + pj
        mask_l = pm * molsize ** 2 + pj * molsize + pi
        out = [nmol, molsize, nSuperHeavy, nHeavy, nHydro, nocc, Z, maskd, atom_molid, mask]
        if return_mask_l:
            out.append(mask_l)
        out += [pm, ni, nj, idxi, idxj, xij, rij]
        return tuple(out)


class Pack_Parameters(torch.nn.Module):
    """Per-atom MOPAC parameters gathered by Z (seqm/basics.py:406-448)."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.method = seqm_parameters["method"]
        if self.method not in METHODS:
            raise ValueError(f"pyseqm_amd supports the sp methods {METHODS}; got {self.method}")
        self.learned_list = seqm_parameters.get("learned", [])
        self.parameters = parameterlist[self.method]
        self.tab = param_table(self.method, seqm_parameters.get("parameter_file_dir"))

    def forward(self, Z, learned_params=dict()):
        dev = Z.device
        zc = Z.detach().cpu().numpy()
        out = dict(learned_params)
        for name in self.parameters:
            if name not in out:
                out[name] = torch.as_tensor(self.tab[zc, COLUMNS.index(name)], dtype=torch.float64, device=dev)
        out["beta"] = torch.stack((out["beta_s"], out["beta_p"]), dim=1)
        zeros = torch.zeros_like(out["U_ss"])
        for k in ("zeta_d", "s_orb_exp_tail", "p_orb_exp_tail", "d_orb_exp_tail", "U_dd", "F0SD", "G2SD",
                  "rho_core"):
            out[k] = zeros
        out["Kbeta"] = out.get("Kbeta", None)
        alpha = torch.zeros(int(Z.max().item()) + 1 if len(Z) else 1, int(Z.max().item()) + 1 if len(Z) else 1,
                            dtype=torch.float64, device=dev)
        return out, alpha, alpha.clone()


def _converger(seqm_parameters):
    conv = seqm_parameters.get("scf_converger", [2])
    kind = int(conv[0])
    if kind == 0:
        return (0, float(conv[1]) if len(conv) > 1 else 0.5)
    if kind in (1, 2):
        return (kind,)
    raise NotImplementedError("scf_converger [3] (KSA) and thermal convergers are outside this path")


class Hamiltonian(torch.nn.Module):
    """SCF on the device (seqm/basics.py:451-533); returns the reference's 12-tuple."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.eps = float(seqm_parameters["scf_eps"])
        self.sp2 = seqm_parameters.get("sp2", [False])
        self.converger = _converger(seqm_parameters)
        self.eig = seqm_parameters.get("eig", True)
        self.scf_backward = seqm_parameters.get("scf_backward", 0)  # basics.py:489-491
        self.scf_backward_eps = float(seqm_parameters.get("scf_backward_eps", seqm_parameters["scf_eps"]))

    def forward(self, molecule, themethod, P0=None):
        from .seqm_functions.scf_loop import scf_loop
        return scf_loop(molecule, eps=self.eps, P=P0, sp2=self.sp2, scf_converger=list(self.converger),
                        eig=self.eig, scf_backward=self.scf_backward, scf_backward_eps=self.scf_backward_eps)


class Energy(torch.nn.Module):
    """Energies of a converged single point (seqm/basics.py:536-1244, ground-state branch)."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.seqm_parameters = seqm_parameters
        self.method = seqm_parameters["method"]
        self.hamiltonian = Hamiltonian(seqm_parameters)
        self.Hf_flag = seqm_parameters.get("Hf_flag", True)

    def forward(self, molecule, learned_parameters=dict(), all_terms=False, P0=None, *args, **kwargs):
        # the reference re-packs the parameters with this call's learned values (basics.py:781-789)
        molecule.use_parameters(learned_parameters)
        eng = molecule.engine()
        res = _run_scf(molecule, self.hamiltonian, P0)
        en = res["energies"]
        trained = _trained_parameters(molecule.learned_parameters)
        if not all_terms:
            if trained:
                raise NotImplementedError("parameter gradients of the per-pair core repulsion EnucAB "
                                          "(all_terms=False) are not supported; use all_terms=True")
            return en["Eelec"], eng_pair_enuc(eng), res["Ppad"], res["notconverged"]
        Hf = en["Hf"] if self.Hf_flag else en["Etot"] - en["Eiso"]
        Etot, Eelec, Enuc, Eiso = en["Etot"], en["Eelec"], en["Enuc"], en["Eiso"]
        if trained:
            names = [n for n, _ in trained]
            Hf, Etot, Eelec, Enuc, Eiso = _ParamEnergies.apply(eng, eng._last_P, names, Hf, Etot, Eelec, Enuc, Eiso,
                                                               *[t for _, t in trained])
        return (Hf, Etot, Eelec, Enuc, Eiso, eng_pair_enuc(eng), res["e_gap"], res["e"],
                res["Ppad"], res["charge"], res["notconverged"])


def _trained_parameters(learned):
    """(name, tensor) of the learned per-atom parameters that ask for a gradient (grad mode on)."""
    if not torch.is_grad_enabled():
        return []
    out = []
    for name, v in (learned or {}).items():
        if torch.is_tensor(v) and v.requires_grad:
            if name not in COLUMNS:
                raise NotImplementedError(f"gradients with respect to {name!r} are not supported "
                                          f"(per-atom parameters: {', '.join(COLUMNS)})")
            out.append((name, v))
    return out


class _ParamEnergies(torch.autograd.Function):
    """Hf, Etot, Eelec, Enuc, Eiso of a converged SCF as functions of the learned per-atom
    parameters.  The backward is the fixed-density derivative (SCFEngine.param_grad), which for the
    energy at a converged SCF density is the full derivative -- the gradient the reference's
    autograd returns with scf_backward = 0 (scf_loop.py:1996-2031, basics.py:781-789); Eiso is
    linear in U and the one-center integrals (energy.py:8-23)."""

    @staticmethod
    def forward(ctx, eng, P, names, Hf, Etot, Eelec, Enuc, Eiso, *params):
        # the engine may be re-parametrised or moved before backward runs: keep what it needs
        ctx.eng, ctx.names = eng, names
        ctx.state = (P.detach().clone(), eng.atom_par.clone(), eng.coords.clone())
        ctx.meta = [(p.shape, p.dtype, p.device) for p in params]
        return Hf.clone(), Etot.clone(), Eelec.clone(), Enuc.clone(), Eiso.clone()

    @staticmethod
    def backward(ctx, gHf, gEtot, gEelec, gEnuc, gEiso):
        eng = ctx.eng
        P, atom_par, coords = ctx.state
        G = eng.param_grad(P, atom_par=atom_par, coords=coords)  # (natoms, 2, 24)
        dI = eng.eiso_grad()
        mol = torch.as_tensor(eng.b.atom_mol, device=G.device).long()
        dT = G[:, 0] + G[:, 1]
        tot = torch.zeros_like(dT)
        for g, d in ((gEelec, G[:, 0]), (gEnuc, G[:, 1]), (gEtot, dT), (gEiso, dI), (gHf, dT - dI)):
            if g is not None:
                tot = tot + g.to(d)[mol][:, None] * d
        grads = [tot[:, COLUMNS.index(n)].reshape(shape).to(dtype=dtype, device=dev)
                 for n, (shape, dtype, dev) in zip(ctx.names, ctx.meta)]
        return (None, None, None, None, None, None, None, None, *grads)


def eng_pair_enuc(eng):
    """Per-pair core-core repulsion in reference pair order."""
    eng.set_coordinates(eng.coords, enuc_pair=True)
    return eng.enuc_pair[:eng.b.npairs].clone()


def _run_scf(molecule, ham, P0):
    from .seqm_functions.scf_loop import scf_loop
    F, e, P, Hcore, w, charge, rho0xi, rho0xj, riXH, ri, notconverged, v = scf_loop(
        molecule, eps=ham.eps, P=P0, sp2=ham.sp2, scf_converger=list(ham.converger), eig=ham.eig,
        scf_backward=ham.scf_backward, scf_backward_eps=ham.scf_backward_eps)
    eng = molecule.engine()
    en = eng.energies(eng._last_P)
    e_gap = None
    if e is not None:
        lumo = molecule.nocc.reshape(-1, 1)
        e_gap = (e.gather(1, lumo) - e.gather(1, lumo - 1)).reshape(-1)
    molecule.w = w
    return dict(energies=en, Ppad=P, e=e, e_gap=e_gap, charge=charge, notconverged=notconverged, v=v)


class Force(torch.nn.Module):
    """Forces -dE/dR (eV/Angstrom) with energies (seqm/basics.py:1247-1365)."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.energy = Energy(seqm_parameters)
        self.seqm_parameters = seqm_parameters

    def forward(self, molecule, learned_parameters=dict(), P0=None, cis_amp=None, do_force=True, *args, **kwargs):
        Hf, Etot, Eelec, Enuc, Eiso, _, e_gap, e, D, charge, notconverged = self.energy(
            molecule, learned_parameters=learned_parameters, all_terms=True, P0=P0)
        eng = molecule.engine()
        if do_force:
            force = eng.b.atom_to_padded(eng.forces(eng._last_P))
            # the reference returns every output detached when it computes forces, and Eelec / Enuc
            # with their parameter gradients otherwise (basics.py:1296-1365)
            Eelec, Enuc = Eelec.detach(), Enuc.detach()
        else:
            force = torch.tensor([])
        return force, D, Hf.detach(), Etot.detach(), Eelec, Enuc, Eiso.detach(), e, e_gap, charge, notconverged
