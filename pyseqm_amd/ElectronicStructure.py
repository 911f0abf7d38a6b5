"""Electronic_Structure: drop-in driver (seqm/ElectronicStructure.py:40-128)."""
import torch

from .basics import Force


class Electronic_Structure(torch.nn.Module):
    def __init__(self, seqm_parameters, *args, **kwargs):
        super().__init__()
        self.seqm_parameters = seqm_parameters
        self.conservative_force = Force(self.seqm_parameters)

    @staticmethod
    def atomic_charges(P, n_orbital=4):
        n_molecule = P.shape[0]
        n_atom = P.shape[1] // n_orbital
        return P.diagonal(dim1=1, dim2=2).reshape(n_molecule, n_atom, n_orbital).sum(axis=2)

    def forward(self, molecule, learned_parameters=dict(), xl_bomd_params=dict(), P0=None, err_threshold=None,
                max_rank=None, T_el=None, dm_prop="SCF", *args, **kwargs):
        if dm_prop == "SCF":
            (molecule.force, P, molecule.Hf, molecule.Etot, molecule.Eelec, molecule.Enuc, molecule.Eiso,
             molecule.e_mo, molecule.e_gap, self.charge, self.notconverged) = self.conservative_force(
                molecule, P0=P0, learned_parameters=learned_parameters)
            molecule.dm = P.detach()
        elif dm_prop == "XL-BOMD":
            from .xlbomd import ForceXL
            (molecule.force, molecule.dm, molecule.Hf, molecule.Etot, molecule.Eelec, molecule.Enuc, molecule.Eiso,
             molecule.e_mo, molecule.e_gap) = ForceXL(self.seqm_parameters)(molecule, P0)
        else:
            raise ValueError(f"dm_prop must be 'SCF' or 'XL-BOMD', got {dm_prop!r}")
        with torch.no_grad():
            molecule.q = molecule.const.tore[molecule.species] - self.atomic_charges(molecule.dm)

    def get_force(self):
        return self.conservative_force
