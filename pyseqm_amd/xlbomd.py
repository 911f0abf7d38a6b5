"""XL-BOMD on the device: density propagation, the XL energy and forces for a packed batch.

One XL-BOMD step (MolecularDynamics.py:1437-1515 with EnergyXL/ForceXL, xlbomd.py:73-551) is

    v += dt/2 a;  x += dt v                                  (k_verlet phase 0)
    P <- kappa (0.95 D + 0.05 P) + sum_j coef[cindx + j] Pt[j];  Pt[m-1-cindx] = P
                                                             (k_xl_propagate)
    integrals + H(x);  F = H + G(P)                          (pair kernels, Fock build)
    D = density(F)                                           (FP64 eigensolve or SP2)
    E = sum D F - 1/2 sum (F - H) P + E_nuc                  (k_xl_energy)
    f = -dE/dx at fixed D, P                                  (k_pair_forces<XL>)
    a = f / m * ACC_SCALE;  v += dt/2 a                       (k_verlet phase 1)

with no SCF loop and no host synchronisation: every array of the trajectory batch stays in HBM.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat

# Niklasson et al. JCP 130, 214109 (2009): kappa, alpha, c0..ck  (MolecularDynamics.py:1382-1390)
XL_COEFFS = {
    3: [1.69, 150e-3, -2.0, 3.0, 0.0, -1.0],
    4: [1.75, 57e-3, -3.0, 6.0, -2.0, -2.0, 1.0],
    5: [1.82, 18e-3, -6.0, 14.0, -8.0, -3.0, 4.0, -1.0],
    6: [1.84, 5.5e-3, -14.0, 36.0, -27.0, -2.0, 12.0, -6.0, 1.0],
    7: [1.86, 1.6e-3, -36.0, 99.0, -88.0, 11.0, 32.0, -25.0, 8.0, -1.0],
    8: [1.88, 0.44e-3, -99.0, 286.0, -286.0, 78.0, 78.0, -90.0, 42.0, -10.0, 1.0],
    9: [1.89, 0.12e-3, -286.0, 858.0, -936.0, 364.0, 168.0, -300.0, 184.0, -63.0, 12.0, -1.0],
}
MIX_C = 0.95  # D/P mixing of the propagation (MolecularDynamics.py:1422)

# unit conversions of the integrator (MolecularDynamics.py:24-31)
ACC_SCALE = 0.009648532800137615      # eV/A / amu -> A/fs^2
VEL_SCALE = 0.9118367323190634e-3     # sqrt(K/amu) -> A/fs
KINETIC_ENERGY_SCALE = 1.0364270099032438e2  # amu (A/fs)^2 -> eV
TEMPERATURE_SCALE = 1.160451812e4     # K/eV


def xl_coefficients(k):
    """(coeff_D, m, coeff[2m]) of the propagation P(n+1) = coeff_D D + sum_j coeff[cindx+j] Pt[j]:
    P(n+1) = 2P(n) - P(n-1) + kappa (D(n) - P(n)) + alpha sum_i c_i P(n-i), regrouped
    (MolecularDynamics.py:1392-1411)."""
    if k not in XL_COEFFS:
        raise ValueError(f"xl_bomd_params['k'] must be one of {sorted(XL_COEFFS)}, got {k}")
    kappa, alpha = XL_COEFFS[k][0], XL_COEFFS[k][1]
    c = np.asarray(XL_COEFFS[k][2:], dtype=np.float64) * alpha
    c[0] += 2.0 - kappa
    c[1] -= 1.0
    return kappa, k + 1, np.concatenate([c, c])


class XLTrajectories:
    """Device-resident XL-BOMD state of a batch of trajectories (one per molecule of the engine).

    Packed arrays: x, v, a, f (natoms, 3) Angstrom / (A/fs) / (A/fs^2) / (eV/A); P, D (total,);
    Pt (m, total) history ring.  ``P0`` is the converged SCF density that seeds P, Pt and D
    (MolecularDynamics.py:1590-1605)."""

    def __init__(self, eng, k, dt, x, v, P0, density="eigh", force_h=1.0e-3, ksa=None):
        """ksa: None, or the max_rank branch of EnergyXL (xlbomd.py:201-339) as a dict with
        "max_rank", "err_threshold", "T_el" and "propagate" (True: P is propagated with the Krylov
        kernel as KSA_XL_BOMD does, MolecularDynamics.py:1615-1619; False: with the density as XL_BOMD
        does).  D is then the Fermi density of F(P) at T_el."""
        self.eng = eng
        b = eng.b
        dev = eng.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.k = int(k)
        self.coeff_D, self.m, coeff = xl_coefficients(self.k)
        self.coeff = [float(c) for c in coeff]
        self.dt = float(dt)
        self.density_kind = density
        self.force_h = float(force_h)
        self.x = torch.as_tensor(x, **f64).reshape(-1, 3).contiguous().clone()
        self.v = torch.as_tensor(v, **f64).reshape(-1, 3).contiguous().clone()
        from .seqm_functions.constants import MASS
        self.minv = torch.as_tensor(1.0 / np.take(MASS, b.Z), **f64)
        self.mass = torch.as_tensor(np.take(MASS, b.Z), **f64)
        total = max(b.total, 1)
        self.P = torch.as_tensor(P0, **f64).reshape(-1).clone()
        self.D = self.P.clone()
        self.Pt = self.P.unsqueeze(0).repeat(self.m, 1).contiguous()
        self.F = torch.empty(total, **f64)
        self.f = torch.zeros(max(b.natoms, 1), 3, **f64)
        self.a = torch.zeros_like(self.f)
        self.eelec = torch.zeros(b.nmol, **f64)
        self.mol_of_atom = torch.as_tensor(b.atom_mol, device=dev).long()
        self.idx_all = torch.arange(b.nmol, dtype=torch.int32, device=dev)
        self.step_count = 0
        self.ksa = dict(ksa) if ksa else None
        self.EEnt = torch.zeros(b.nmol, **f64)  # -2 T_el S, the XL potential's entropy term
        if self.ksa:
            from .ksa import KSAWorkspace
            self.ks = KSAWorkspace(eng, int(self.ksa["max_rank"]))
            self.dP2dt2 = torch.zeros(total, **f64)
            self.tmp = torch.zeros(total, **f64)
            self.STATE = XLTrajectories.STATE + ("dP2dt2", "EEnt")

    # ----------------------------------------------------------------- pieces
    def set_forces(self, f):
        """Accelerations from forces (natoms, 3) eV/A (MolecularDynamics.py:903-904)."""
        self.f.copy_(f)
        self.a.copy_(f * self.minv[:, None] * ACC_SCALE)

    def _verlet(self, phase):
        b = self.eng.b
        nat.call_raw("pyseqm_md_verlet", b.natoms, phase, self.dt, ACC_SCALE, nat.ptr(self.x), nat.ptr(self.v),
                     nat.ptr(self.a), nat.ptr(self.f), nat.ptr(self.minv), self.eng.stream)

    def _propagate(self):
        cindx = self.step_count % self.m
        coef = (ctypes.c_double * self.m)(*self.coeff[cindx:cindx + self.m])
        D, c = self.D, MIX_C
        if self.ksa and self.ksa["propagate"]:
            # KSA_XL_BOMD._propagate_P: kappa (dP2dt2 + P) + the history terms
            from .ksa import _axpy
            self.tmp.copy_(self.P)
            _axpy(self.eng, self.ks, 4, 1.0, None, 0, self.dP2dt2, self.tmp)
            D, c = self.tmp, 1.0
        nat.call_raw("pyseqm_xl_propagate", self.eng.b.total, nat.ptr(D), nat.ptr(self.P), nat.ptr(self.Pt),
                     self.Pt.shape[1], self.m, coef, self.coeff_D, c, self.m - 1 - cindx, self.eng.stream)

    def energy_and_forces(self):
        """At the current x and P: F(P), D = density(F), the XL energy and forces at fixed D, P."""
        eng, b = self.eng, self.eng.b
        eng.set_coordinates(self.x)
        eng.fock(self.P, self.F)
        if self.ksa:
            from . import ksa as K
            T = float(self.ksa["T_el"])
            K.fermi_density(eng, self.ks, self.F, T, self.D)
            torch.mul(self.ks.S, -2.0 * T, out=self.EEnt)
            K.rank_kernel(eng, self.ks, self.D, self.P, T, float(self.ksa.get("err_threshold", 0.0)), self.dP2dt2)
        else:
            eng.density_kind = self.density_kind
            eng._idx_host = np.arange(b.nmol)
            eng.density(self.F, self.idx_all, b.nmol, self.D)
        nat.call("pyseqm_xl_energy", b.L, nat.ptr(self.D), nat.ptr(self.P), nat.ptr(self.F), nat.ptr(eng.H),
                 nat.ptr(self.eelec), nat.ptr(eng.red), nat.red_blocks(b.nmax), eng.stream)
        nat.call("pyseqm_pair_forces_xl", b.L, nat.METHODS[eng.method], nat.ptr(self.x), nat.ptr(eng.atom_par),
                 nat.ptr(self.D), nat.ptr(self.P), self.force_h, nat.ptr(eng.part), nat.ptr(self.f), eng.stream)

    def step(self):
        """One XL-BOMD step of every trajectory (MolecularDynamics.py:1437-1515), device only."""
        self._verlet(0)
        self._propagate()
        self.energy_and_forces()
        self._verlet(1)
        self.step_count += 1

    # ---------------------------------------------------------- checkpoints
    STATE = ("x", "v", "a", "f", "P", "D", "Pt", "eelec")

    def state_dict(self):
        """The trajectory batch's device state (positions, velocities, accelerations, forces, the
        density field P, D, the history ring Pt, the XL energies) and its step counter: what a run
        needs to continue bit for bit (the integrals and Fock matrices follow from x and P)."""
        return {"tensors": {k: getattr(self, k) for k in self.STATE},
                "meta": {"k": self.k, "m": self.m, "dt": self.dt, "step_count": self.step_count,
                         "density": self.density_kind, "force_h": self.force_h, "ksa": self.ksa,
                         "natoms": int(self.eng.b.natoms), "total": int(self.eng.b.total)}}

    def load_state_dict(self, state):
        """Copy a state_dict into this batch's device tensors, in place (graphs captured on them stay
        valid); the batch must hold the same molecules, k, time step and density step."""
        meta, mine = state["meta"], self.state_dict()["meta"]
        for key in ("k", "m", "natoms", "total", "density", "ksa"):
            if meta[key] != mine[key]:
                raise ValueError(f"checkpoint {key} = {meta[key]!r} does not match this trajectory batch ({mine[key]!r})")
        if float(meta["dt"]) != self.dt:
            raise ValueError(f"checkpoint time step {meta['dt']} != {self.dt}")
        with torch.no_grad():
            for k in self.STATE:
                src, dst = state["tensors"][k], getattr(self, k)
                if tuple(src.shape) != tuple(dst.shape):
                    raise ValueError(f"checkpoint {k}: shape {tuple(src.shape)} != {tuple(dst.shape)}")
                dst.copy_(src.to(dst.device, dst.dtype))
        self.step_count = int(meta["step_count"])

    # ------------------------------------------------------------- thermo
    def energies(self):
        eng = self.eng
        etot = self.eelec + eng.enuc
        return dict(Eelec=self.eelec.clone(), Enuc=eng.enuc.clone(), Etot=etot, Eiso=eng.eiso.clone(),
                    Hf=etot - eng.eiso + eng.eheat)

    def kinetic_energy(self):
        """Per-molecule 1/2 sum m v^2 in eV (MolecularDynamics.py:784-789)."""
        e = 0.5 * self.mass * (self.v * self.v).sum(1) * KINETIC_ENERGY_SCALE
        out = torch.zeros(self.eng.b.nmol, dtype=torch.float64, device=self.eng.device)
        return out.index_add_(0, self.mol_of_atom, e)


class EnergyXL(torch.nn.Module):
    """Energies of the XL Lagrangian at a given density field P (xlbomd.py:73-508, ground state):
    returns (Hf, Etot, Eelec, EEnt, Enuc, Eiso, EnucAB, D, dP2dt2, Error, e_gap, e, Fe_occ) with D the
    density of F(P), or with xl_bomd_params["max_rank"] the Fermi density at T_el, its entropy term
    EEnt = -2 T_el S and the rank-m Krylov kernel dP2dt2 (xlbomd.py:201-341)."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.seqm_parameters = seqm_parameters
        self.sp2 = seqm_parameters.get("sp2", [False])

    def _run(self, molecule, P, xl_bomd_params):
        eng = molecule.engine()
        b = eng.b
        ksa = None
        if "max_rank" in xl_bomd_params:
            ksa = {"max_rank": int(xl_bomd_params["max_rank"]),
                   "err_threshold": float(xl_bomd_params.get("err_threshold", 0.0)),
                   "T_el": float(xl_bomd_params["T_el"]), "propagate": False}
        tr = getattr(molecule, "_xl", None)
        Pp = b.from_padded(P.detach().to(torch.float64))
        x = b.atoms_from_padded(molecule.coordinates.detach())
        if tr is None or tr.ksa != ksa:
            tr = XLTrajectories(eng, 3, 0.0, x, torch.zeros_like(x), Pp,
                                density="sp2" if self.sp2[0] else "eigh", ksa=ksa)
            molecule._xl = tr
        tr.x.copy_(b.atoms_from_padded(x))
        tr.P.copy_(Pp)
        tr.energy_and_forces()
        return eng, tr

    def forward(self, molecule, P, cis_amp=None, learned_parameters=dict(), xl_bomd_params=dict(), all_terms=False,
                *args, **kwargs):
        eng, tr = self._run(molecule, P, xl_bomd_params)
        en = tr.energies()
        D = eng.b.to_padded(tr.D)
        if tr.ksa:
            b, ks = eng.b, tr.ks
            e = torch.zeros(b.nmol, int(b.norb.max()), dtype=torch.float64, device=eng.device)
            occ = torch.zeros_like(e)
            offs = np.concatenate([[0], np.cumsum(b.norb)[:-1]])
            for m in range(b.nmol):  # eigenvalues / occupations, one padded row per molecule
                o, n = int(offs[m]), int(b.norb[m])
                e[m, :n] = ks.evals[o:o + n]
                occ[m, :n] = ks.occ[o:o + n]
            dP2dt2 = b.to_padded(tr.dP2dt2)
            err = ks.last_error
            fe_occ = occ
        else:
            e, _ = eng.eig(tr.F)
            dP2dt2 = err = fe_occ = None
        lumo = molecule.nocc.reshape(-1, 1)
        e_gap = (e.gather(1, lumo) - e.gather(1, lumo - 1)).reshape(-1)
        ent = tr.EEnt.clone()
        if not all_terms:
            return en["Eelec"], None, D, e_gap, e
        return (en["Hf"], en["Etot"], en["Eelec"], ent, en["Enuc"], en["Eiso"], None, D, dP2dt2, err, e_gap, e,
                fe_occ)


class ForceXL(torch.nn.Module):
    """XL-BOMD forces (xlbomd.py:511-569): -d Hf/dR at fixed D and P.  Returns (force, D, Hf, Etot,
    Eelec, Enuc, Eiso, e, e_gap) in the reference's padded layouts."""

    def __init__(self, seqm_parameters):
        super().__init__()
        self.energy = EnergyXL(seqm_parameters)
        self.seqm_parameters = seqm_parameters

    def forward(self, molecule, P, cis_amp=None, learned_parameters=dict(), xl_bomd_params=dict(), *args, **kwargs):
        Hf, Etot, Eelec, EEnt, Enuc, Eiso, _, D, _, _, e_gap, e, _ = self.energy(
            molecule, P, cis_amp, learned_parameters, xl_bomd_params=xl_bomd_params, all_terms=True)
        tr = molecule._xl
        force = tr.eng.b.atom_to_padded(tr.f[:tr.eng.b.natoms].clone())
        molecule.Electronic_entropy = EEnt
        return force, D, Hf, Etot, Eelec, Enuc, Eiso, e, e_gap


def save_state(path, state, extra=None):
    """Write a state_dict (plus JSON-serialisable `extra`) to `path` as safetensors, atomically (a
    temporary file renamed over the target); reading it back executes nothing from the file."""
    import json
    import os

    from safetensors.torch import save_file
    meta = {"meta": json.dumps(state["meta"]), "extra": json.dumps(extra or {})}
    tmp = f"{path}.tmp{os.getpid()}"
    save_file({k: v.detach().contiguous().cpu() for k, v in state["tensors"].items()}, tmp, metadata=meta)
    os.replace(tmp, path)


def load_state(path):
    """(state_dict with host tensors, extra) from a file written by save_state."""
    import json

    from safetensors import safe_open
    with safe_open(path, framework="pt", device="cpu") as fh:
        md = fh.metadata()
        tensors = {k: fh.get_tensor(k) for k in fh.keys()}
    return {"tensors": tensors, "meta": json.loads(md["meta"])}, json.loads(md["extra"])
