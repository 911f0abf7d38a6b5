"""Molecular dynamics drivers of the drop-in API (seqm/MolecularDynamics.py).

``Molecular_Dynamics_Basic`` (velocity Verlet on converged SCF forces, 670-1092) and ``XL_BOMD``
(extended-Lagrangian Born-Oppenheimer MD, 1373-1605) keep the reference's constructor and
``run(molecule, steps, ..., seed)`` signatures.  The XL-BOMD trajectory batch lives on the device
(``XLTrajectories``): one step is a fixed sequence of kernel launches with no host round trip.
Thermodynamic records (T, Ek, Ep per step) are kept in memory in ``self.history``; the HDF5/xyz
writers of the reference are not part of this path.  ``XL_BOMD.save_checkpoint`` /
``XL_BOMD.run_from_checkpoint`` (MolecularDynamics.py:1099-1166) write and resume the device-resident
trajectory state (safetensors, nothing executed on load), continuing a run bit for bit.
"""
import warnings

import numpy as np
import torch

from .ElectronicStructure import Electronic_Structure
from .xlbomd import (KINETIC_ENERGY_SCALE, TEMPERATURE_SCALE, VEL_SCALE, ACC_SCALE, XLTrajectories)


class Molecular_Dynamics_Basic(torch.nn.Module):
    def __init__(self, seqm_parameters, timestep=1.0, Temp=0.0, step_offset=0, output=None, *args, **kwargs):
        super().__init__()
        self.seqm_parameters = seqm_parameters
        self.timestep = float(timestep)
        self.Temp = float(Temp)
        self.step_offset = int(step_offset)
        self.output = output
        self.esdriver = Electronic_Structure(seqm_parameters)
        self.n_dof = None
        self.history = None

    # -------------------------------------------------------- thermodynamics
    def _kinetic_energy(self, molecule):
        return torch.sum(0.5 * molecule.mass * molecule.velocities ** 2, dim=(1, 2)) * KINETIC_ENERGY_SCALE

    def _calc_temperature(self, Ek):
        return Ek * TEMPERATURE_SCALE / (0.5 * self.n_dof)

    def set_dof(self, molecule, constraints=0.0):
        self.n_dof = 3.0 * molecule.num_atoms - constraints

    def _thermo_potential(self, molecule):
        return molecule.Etot

    def initialize_velocity(self, molecule, vel_com=True):
        """Maxwell-Boltzmann velocities rescaled to exactly Temp (MolecularDynamics.py:717-744).
        The normal deviates come from the generator of the molecule's device, in the padded
        layout (torch.randn_like of the coordinates, seeded by ``run(seed=)``), as in the
        reference."""
        if torch.is_tensor(molecule.velocities):
            if vel_com:
                self._zero_com(molecule, translate_to_origin=False, restore_kinetic_energy=False)
            return molecule.velocities
        if self.Temp == 0.0:
            molecule.velocities = torch.zeros_like(molecule.coordinates)
            return molecule.velocities
        scale = torch.sqrt(self.Temp * molecule.mass_inverse) * VEL_SCALE
        molecule.velocities = torch.randn_like(molecule.coordinates.detach()) * scale
        Ek = self._kinetic_energy(molecule)
        alpha = torch.sqrt(self.Temp / self._calc_temperature(Ek))
        molecule.velocities.mul_(alpha.reshape(-1, 1, 1))
        if vel_com:
            self._zero_com(molecule, translate_to_origin=True)
        return molecule.velocities

    def _zero_com(self, molecule, remove_angular=True, translate_to_origin=False, restore_kinetic_energy=True):
        """Remove centre-of-mass translation and rotation (MolecularDynamics.py:746-782)."""
        mass = molecule.mass
        M = mass.sum(dim=1, keepdim=True)
        ek0 = self._kinetic_energy(molecule)
        with torch.no_grad():
            x = molecule.coordinates
            r = x - (mass * x).sum(dim=1, keepdim=True) / M
            if translate_to_origin:
                molecule.coordinates.copy_(r)
            v = molecule.velocities
            v.sub_((mass * v).sum(dim=1, keepdim=True) / M)
            if remove_angular:
                Lmom = (mass * torch.linalg.cross(r, v, dim=2)).sum(dim=1)
                rr = (r * r).sum(dim=2, keepdim=True)
                eye = torch.eye(3, dtype=x.dtype, device=x.device)
                inertia = (mass * rr).sum(dim=1, keepdim=True) * eye - (mass.unsqueeze(3) * r.unsqueeze(3)
                                                                          * r.unsqueeze(2)).sum(dim=1)
                omega = (torch.linalg.pinv(inertia, hermitian=True, atol=1e-10) @ Lmom.unsqueeze(2)).squeeze(-1)
                v.sub_(torch.linalg.cross(omega.unsqueeze(1).expand_as(r), r, dim=2))
            ek1 = self._kinetic_energy(molecule)
            if torch.any(ek1 < 1e-12):
                raise RuntimeError("Zero kinetic energy after removing COM momentum")
            if restore_kinetic_energy:
                v.mul_(torch.sqrt(ek0 / ek1).reshape(-1, 1, 1))

    # ------------------------------------------------------------- stepping
    def one_step(self, molecule, learned_parameters=dict(), *args, **kwargs):
        """Velocity Verlet on SCF forces (MolecularDynamics.py:815-842)."""
        dt = self.timestep
        with torch.no_grad():
            molecule.velocities.add_(0.5 * molecule.acc * dt)
            molecule.coordinates.add_(molecule.velocities * dt)
        self.esdriver(molecule, learned_parameters=learned_parameters, P0=molecule.dm, dm_prop="SCF")
        with torch.no_grad():
            molecule.acc = molecule.force * molecule.mass_inverse * ACC_SCALE
            molecule.velocities.add_(0.5 * molecule.acc * dt)

    def _do_integrator_step(self, i, molecule, learned_parameters, **kwargs):
        return self.one_step(molecule, learned_parameters=learned_parameters, **kwargs)

    def initialize(self, molecule, remove_com=None, learned_parameters=dict(), steps=None, *args, **kwargs):
        if remove_com is not None:
            raise NotImplementedError("periodic COM removal during the run is not part of this path")
        self.set_dof(molecule, 0.0)
        if self.step_offset == 0 or not torch.is_tensor(molecule.velocities):
            self.initialize_velocity(molecule)
        if not torch.is_tensor(molecule.force):
            self.esdriver(molecule, learned_parameters=learned_parameters, P0=molecule.dm)
        with torch.no_grad():
            molecule.acc = molecule.force * molecule.mass_inverse * ACC_SCALE
        self.history = {"T": [], "Ek": [], "Ep": []}
        if self.step_offset == 0:
            self._record(molecule)

    def _record(self, molecule):
        with torch.no_grad():
            Ek = self._kinetic_energy(molecule)
            self.history["Ek"].append(Ek.detach().cpu().numpy())
            self.history["T"].append(self._calc_temperature(Ek).detach().cpu().numpy())
            self.history["Ep"].append(self._thermo_potential(molecule).detach().cpu().numpy())

    def run(self, molecule, steps, learned_parameters=dict(), reuse_P=True, remove_com=None, seed=None, *args,
            **kwargs):
        if seed is not None:
            torch.manual_seed(int(seed))
            if torch.cuda.is_available():
                torch.cuda.manual_seed_all(int(seed))
        self.initialize(molecule, remove_com=remove_com, learned_parameters=learned_parameters, steps=steps)
        if not reuse_P:
            molecule.dm = None
        for i in range(self.step_offset, steps):
            self._do_integrator_step(i, molecule, learned_parameters, **kwargs)
            if not reuse_P:
                molecule.dm = None
            self._record(molecule)
        return molecule.coordinates, molecule.velocities, molecule.acc

    def thermo(self):
        """(T, Ek, Ep, Etot) arrays of shape (nsteps + 1, nmol)."""
        T, Ek, Ep = (np.array(self.history[k]) for k in ("T", "Ek", "Ep"))
        return T, Ek, Ep, Ek + Ep


class XL_BOMD(Molecular_Dynamics_Basic):
    """XL-BOMD (MolecularDynamics.py:1373-1605) with the trajectory batch resident on the device."""

    KSA_PROPAGATION = False  # KSA_XL_BOMD propagates P with the Krylov kernel dP2dt2

    def __init__(self, damp=None, xl_bomd_params=dict(), *args, **kwargs):
        if damp is not None:
            raise NotImplementedError("the Langevin-damped XL-BOMD variant is not part of this path")
        self.k = int(xl_bomd_params["k"])
        self.xl_bomd_params = xl_bomd_params
        self.damp = damp
        super().__init__(*args, **kwargs)
        self.traj = None

    def _thermo_potential(self, molecule):
        return molecule.Etot + molecule.Electronic_entropy

    def initialize(self, molecule, remove_com=None, learned_parameters=dict(), steps=None, *args, **kwargs):
        molecule.Electronic_entropy = torch.zeros(molecule.species.shape[0], dtype=torch.float64,
                                                  device=molecule.coordinates.device)
        super().initialize(molecule, remove_com=remove_com, learned_parameters=learned_parameters, steps=steps)
        eng = molecule.engine()
        b = eng.b
        sp2 = self.seqm_parameters.get("sp2", [False])
        self.traj = XLTrajectories(eng, self.k, self.timestep, b.atoms_from_padded(molecule.coordinates.detach()),
                                   b.atoms_from_padded(molecule.velocities.detach()), b.from_padded(molecule.dm),
                                   density="sp2" if sp2[0] else "eigh", ksa=self._ksa_params())
        self.traj.set_forces(b.atoms_from_padded(molecule.force.detach()))

    def _ksa_params(self):
        """The max_rank branch of EnergyXL (xlbomd.py:201-341): the Fermi density at T_el and the
        rank-m Krylov kernel; KSA_XL_BOMD also propagates with it."""
        p = self.xl_bomd_params
        if "max_rank" not in p:
            if self.KSA_PROPAGATION:
                raise ValueError("KSA_XL_BOMD needs xl_bomd_params['max_rank'] (and 'T_el', 'err_threshold')")
            return None
        if self.seqm_parameters.get("sp2", [False])[0]:
            raise NotImplementedError("the Fermi density of the max_rank branch comes from the eigen-decomposition; "
                                      "sp2 is not used with it")
        return {"max_rank": int(p["max_rank"]), "err_threshold": float(p.get("err_threshold", 0.0)),
                "T_el": float(p["T_el"]), "propagate": bool(self.KSA_PROPAGATION)}

    def _do_integrator_step(self, i, molecule, learned_parameters, **kwargs):
        tr = self.traj
        tr.step_count = i
        tr.step()
        self._sync_molecule(molecule)

    # --------------------------------------------------------- checkpoints
    def save_checkpoint(self, molecule, steps, reuse_P=True, remove_com=None, *, step_done, path):
        """Write the trajectory batch after `step_done` steps of a `steps`-step run (the reference's
        save_checkpoint, MolecularDynamics.py:1099-1116): device state, thermo history, parameters."""
        from .xlbomd import save_state
        extra = {"MD_type": type(self).__name__, "seqm_parameters": self.seqm_parameters,
                 "xl_bomd_params": self.xl_bomd_params,
                 "timestep": self.timestep, "Temp": self.Temp, "steps": int(steps), "step_done": int(step_done),
                 "reuse_P": bool(reuse_P), "species": molecule.species.detach().cpu().tolist(),
                 "charges": molecule.tot_charge.detach().cpu().tolist() if torch.is_tensor(molecule.tot_charge) else 0,
                 "history": {k: [np.asarray(v).tolist() for v in vals] for k, vals in self.history.items()}}
        save_state(path, self.traj.state_dict(), extra)

    @staticmethod
    def run_from_checkpoint(path, device=None):
        """Resume a run written by save_checkpoint and carry it to its `steps` (the reference's
        run_from_checkpoint, MolecularDynamics.py:1124-1166).  Returns (md, molecule)."""
        from .Molecule import Molecule
        from .seqm_functions.constants import Constants
        from .xlbomd import load_state
        state, extra = load_state(path)
        cls = {"XL_BOMD": XL_BOMD, "KSA_XL_BOMD": KSA_XL_BOMD}.get(extra.get("MD_type"))
        if cls is None:
            raise RuntimeError(f"unknown MD type {extra.get('MD_type')!r} in checkpoint")
        dev = torch.device(device) if device is not None else torch.device("cuda")
        species = torch.as_tensor(extra["species"], dtype=torch.int64, device=dev)
        const = Constants().to(dev)
        from .layout import PackedBatch
        b0 = PackedBatch(species.cpu().numpy(), extra["seqm_parameters"]["method"], "cpu")
        x = state["tensors"]["x"][:b0.natoms].numpy()
        coords = np.zeros((species.shape[0], species.shape[1], 3))
        coords.reshape(-1, 3)[b0.real_flat] = x
        charges = torch.as_tensor(extra["charges"], dtype=torch.float64, device=dev) \
            if isinstance(extra["charges"], list) else extra["charges"]
        molecule = Molecule(const, extra["seqm_parameters"], torch.as_tensor(coords, device=dev), species,
                            charges=charges)
        md = cls(xl_bomd_params=extra["xl_bomd_params"], seqm_parameters=extra["seqm_parameters"],
                 timestep=extra["timestep"], Temp=extra["Temp"], step_offset=extra["step_done"])
        md.set_dof(molecule, 0.0)
        eng = molecule.engine()
        bt = eng.b
        sp2 = extra["seqm_parameters"].get("sp2", [False])
        t = state["tensors"]
        md.traj = XLTrajectories(eng, md.k, md.timestep, t["x"].to(dev), t["v"].to(dev), t["P"].to(dev),
                                 density="sp2" if sp2[0] else "eigh", ksa=md._ksa_params())
        md.traj.load_state_dict(state)
        eng.set_coordinates(md.traj.x)
        molecule.Electronic_entropy = torch.zeros(bt.nmol, dtype=torch.float64, device=dev)
        md._sync_molecule(molecule)
        md.history = {k: [np.asarray(v) for v in vals] for k, vals in extra["history"].items()}
        reuse_P = extra["reuse_P"]
        for i in range(md.step_offset, extra["steps"]):
            md._do_integrator_step(i, molecule, dict())
            if not reuse_P:
                molecule.dm = None
            md._record(molecule)
        return md, molecule

    def _sync_molecule(self, molecule):
        """Write the device state back into the reference's padded attributes."""
        tr = self.traj
        b = tr.eng.b
        n = b.natoms
        with torch.no_grad():
            molecule.coordinates.data.copy_(b.atom_to_padded(tr.x[:n]))
            molecule.velocities = b.atom_to_padded(tr.v[:n].clone())
            molecule.acc = b.atom_to_padded(tr.a[:n].clone())
            molecule.force = b.atom_to_padded(tr.f[:n].clone())
            en = tr.energies()
            molecule.Etot, molecule.Eelec, molecule.Enuc = en["Etot"], en["Eelec"], en["Enuc"]
            molecule.Hf, molecule.Eiso = en["Hf"], en["Eiso"]
            molecule.dm = b.to_padded(tr.D)
            molecule.Electronic_entropy = tr.EEnt.clone()
            if tr.ksa:
                molecule.dP2dt2 = b.to_padded(tr.dP2dt2)
                # eigenvectors the Fermi density's eigensolver could not normalise, counted on the
                # device over the steps since the last write-back
                lost = int(tr.ks.fail.item())
                if lost:
                    tr.ks.fail.zero_()
                    warnings.warn(f"KSA-XL-BOMD: the eigensolver lost {lost} eigenvector(s) to a cluster's span "
                                  "since the last write-back; the thermal densities involved may be inaccurate")


class KSA_XL_BOMD(XL_BOMD):
    """Krylov-subspace XL-BOMD (MolecularDynamics.py:1608-1619): the XL step with the Fermi density
    at T_el and P propagated with the rank-m Krylov kernel, P(n+1) = kappa (dP2dt2 + P(n)) + the
    history terms (xl_bomd_params: k, max_rank, err_threshold, T_el).  The potential energy of the
    thermo output includes the electronic entropy term -2 T_el S."""

    KSA_PROPAGATION = True
