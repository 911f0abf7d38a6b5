"""Molecule: the batch container of the drop-in API (seqm/Molecule.py:12-206).

Takes the reference's inputs -- padded, Z-sorted ``species`` (nmol, molsize) and ``coordinates``
(nmol, molsize, 3) in Angstrom -- and exposes the same attributes (parser outputs, parameters,
result slots).  Internally it owns the packed device layout the HIP kernels run on.
"""
from typing import Optional

import numpy as np
import torch

from .basics import Pack_Parameters, Parser
from .layout import check_species


class Molecule(torch.nn.Module):
    def __init__(self, const, seqm_parameters, coordinates, species, charges=0, mult=1,
                 learned_parameters=dict(), do_large_tensors=True, *args, **kwargs):
        super().__init__()
        self.const = const
        check_input(species)
        self.species = species
        self.coordinates = torch.nn.Parameter(coordinates)
        self.coordinates.requires_grad_(False)
        if not torch.is_tensor(charges):
            charges = charges * torch.ones(coordinates.shape[0], device=coordinates.device)
        self.tot_charge = charges
        if not torch.is_tensor(mult):
            mult = mult * torch.ones(coordinates.shape[0], device=coordinates.device)
        self.mult = mult
        if seqm_parameters.get("elements") is None:
            seqm_parameters["elements"] = [0] + sorted(set(species.reshape(-1).tolist()))
        self.seqm_parameters = seqm_parameters
        self.method = seqm_parameters["method"]
        if seqm_parameters.get("UHF", False):
            raise NotImplementedError("pyseqm_amd implements the closed-shell (RHF) path only")
        self.parser = Parser(seqm_parameters)
        self.packpar = Pack_Parameters(seqm_parameters)
        (self.nmol, self.molsize, self.nSuperHeavy, self.nHeavy, self.nHydro, self.nocc, self.Z, self.maskd,
         self.atom_molid, self.mask, self.mask_l, self.pair_molid, self.ni, self.nj, self.idxi, self.idxj,
         self.xij, self.rij) = self.parser(self, self.method, return_mask_l=True)
        if callable(learned_parameters):
            learned_parameters = learned_parameters(self.species, self.coordinates)
        self.learned_parameters = dict(learned_parameters)
        self.parameters, self.alp, self.chi = self.packpar(self.Z, learned_params=dict(learned_parameters))
        self.norb = self.nHydro + 4 * self.nHeavy
        non_zero = self.species != 0
        self.num_atoms = non_zero.sum(dim=1).to(coordinates.dtype)
        self.mass = self.const.mass[self.species].unsqueeze(2)
        self.mass_inverse = torch.zeros_like(self.mass)
        self.mass_inverse[non_zero] = 1.0 / self.mass[non_zero]
        self.force = None
        self.velocities = None
        self.acc = None
        self.dm: Optional[torch.Tensor] = None
        self.q: Optional[torch.Tensor] = None
        self.w: Optional[torch.Tensor] = None
        self.Hf = self.Etot = self.Eelec = self.Enuc = self.Eiso = None
        self.e_mo = self.e_gap = None
        self.molecular_orbitals = None
        self.charge = None
        self.notconverged = None
        self.active_state = 0
        self._engine = None

    # device engine bound to this batch (built on first use, reused across MD steps); it runs with
    # the parameter file of seqm_parameters["parameter_file_dir"] and this molecule's learned values
    def engine(self):
        if self._engine is None:
            from .engine import SCFEngine
            self._engine = SCFEngine(self.species, None, self.method, self.coordinates.device,
                                     charges=self.tot_charge,
                                     param_dir=self.seqm_parameters.get("parameter_file_dir"))
            if self.learned_parameters:
                self._engine.set_parameters(self.parameters)
        return self._engine

    def use_parameters(self, learned_parameters):
        """Re-pack the parameters with `learned_parameters` (a dict or a callable of (species,
        coordinates)) and load them into the engine, as Energy.forward does on every call
        (basics.py:781-789): an empty dict means the parameter file's values."""
        if callable(learned_parameters):
            learned_parameters = learned_parameters(self.species, self.coordinates)
        learned_parameters = dict(learned_parameters)
        eng = self.engine()
        if not learned_parameters and eng.default_params and not self.learned_parameters:
            return
        self.parameters, self.alp, self.chi = self.packpar(self.Z, learned_params=dict(learned_parameters))
        self.learned_parameters = learned_parameters
        eng.set_parameters(self.parameters)
        eng.default_params = not learned_parameters

    def get_coordinates(self):
        return self.coordinates

    def get_species(self):
        return self.species


def check_input(species):
    """Rows of species must be non-increasing (seqm/Molecule.py:188-206)."""
    check_species(species.detach().cpu().numpy() if torch.is_tensor(species) else np.asarray(species))
