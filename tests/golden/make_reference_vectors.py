"""Write tests/golden/reference_vectors.json from the lanl/PYSEQM reference, run on the CPU.

Run in the build container only (the reference is not on the GPU box and does not travel):

    python tests/golden/make_reference_vectors.py [/path/to/PYSEQM]

What it pins that the reference's own goldens do not hold:
  * the constants tables behind the heat of formation (eheat, the isolated-atom coefficients
    ussc..hspc, tore) - seqm/seqm_functions/constants.py;
  * Eiso and Hf = Etot - sum Eiso + sum eheat (seqm/seqm_functions/energy.py:8-23, 195-216) end to end;
  * absolute energies (Hf, Etot, Eelec, Enuc) of the synthetic benchmark inputs: PM3 benzene (as the
    16-benzene batch), an AM1 alkane slice whose longest chains put atom pairs beyond the
    40-bohr overlap cutoff (seqm/seqm_functions/hcore.py:83), and PM3 / MNDO water clusters
    (O-O and O-H pairs, which no reference golden covers).
The reference is imported with an empty stand-in for h5py, which it imports at package import time
but never calls on the single-point path (seqm/MolecularDynamics.py:10).  It runs the energy
module that Force(seqm_parameters)(molecule, do_force=False) runs (seqm/basics.py:1278), in
float64.  Large inputs are stored as the generator call that makes them plus a checksum of the
coordinates (the tests regenerate them and check the checksum); small ones are stored whole.
"""
import contextlib
import io
import json
import os
import re
import sys
import types

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from pyseqm_amd import workloads as W  # noqa: E402
from tests.conftest import read_xyz  # noqa: E402  (before the reference's own tests package is on the path)

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def import_reference():
    h5 = types.ModuleType("h5py")
    h5.__getattr__ = lambda name: type(name, (), {})  # type annotations only (h5py.Group, h5py.File)
    sys.modules.setdefault("h5py", h5)
    sys.path.append(REF)
    from seqm.basics import Force
    from seqm.Molecule import Molecule
    from seqm.seqm_functions.constants import Constants
    return Force, Molecule, Constants


def run(Force, Molecule, Constants, species, coords, method, eps, converger):
    torch.set_default_dtype(torch.float64)
    const = Constants()
    sp = torch.as_tensor(species, dtype=torch.int64)
    xyz = torch.as_tensor(coords, dtype=torch.float64)
    par = {"method": method, "scf_eps": eps, "scf_converger": list(converger), "sp2": [False],
           "pair_outer_cutoff": 1.0e10, "eig": True}
    mol = Molecule(const, par, xyz, sp)
    drv = Force(par)
    log = io.StringIO()
    with contextlib.redirect_stdout(log):
        Hf, Etot, Eelec, Enuc, Eiso, *_rest = drv.energy(mol, all_terms=True)
    # the SCF prints one line per iteration, "scf <mixer> step : <k> | ..." (scf_loop.py:601, 978)
    steps = [int(m) for m in re.findall(r"^scf [a-z ]+:\s*(\d+) \|", log.getvalue(), re.M)]
    return dict(Hf=Hf.detach().tolist(), Etot=Etot.detach().tolist(), Eelec=Eelec.detach().tolist(),
                Enuc=Enuc.detach().tolist(), Eiso=Eiso.detach().tolist(), scf_iterations=max(steps))


def checksum(coords):
    x = np.asarray(coords, dtype=np.float64)
    return [float(x.sum()), float((x * x).sum())]


def case(name, species, coords, method, eps=1.0e-8, converger=(2,), generator=None, **extra):
    c = dict(name=name, method=method, eps=eps, converger=list(converger), _species=species, _coords=coords,
             **extra)
    if generator is None:
        c.update(species=np.asarray(species).tolist(), coordinates=np.asarray(coords).tolist())
    else:
        c.update(generator=generator, coord_checksum=checksum(coords))
    return c


def cases():
    out = []
    # PM3 benzene with the benchmark's noise (two of the 16-benzene batch)
    sp, x = W.benzene(16, seed=0)
    out.append(case("pm3_benzene16_c1", sp, x, "PM3", generator=dict(fn="benzene", args=[16], seed=0)))
    # AM1 alkanes: the longest and shortest chains of the 1024-alkane batch plus some in between
    sp, x = W.alkanes(1024, seed=0)
    nC = (sp == 6).sum(1)
    pick = [int(np.argmax(nC)), int(np.argmin(nC))]
    for target in (9, 17, 23):
        pick.append(int(np.nonzero(nC == target)[0][0]))
    pick = sorted(set(pick))
    K = int((sp[pick] > 0).sum(1).max())
    out.append(case("am1_alkanes_c2_slice", sp[pick][:, :K], x[pick][:, :K], "AM1", batch_index=pick))
    # the first 128 molecules of the 1024-alkane batch (uneven sizes, 14-182 orbitals)
    K = int((sp[:128] > 0).sum(1).max())
    out.append(case("am1_alkanes_c2_first128", sp[:128, :K], x[:128, :K], "AM1",
                    generator=dict(fn="alkanes", args=[1024], seed=0, stop=128)))
    # PM3 and MNDO water clusters built like the (H2O)_256 clusters, smaller
    for nw, method in ((8, "PM3"), (16, "PM3"), (8, "MNDO"), (8, "AM1")):
        sp, x = W.water_clusters(nw, 1, seed=0)
        out.append(case(f"{method.lower()}_water{nw}", sp, x, method))
    sp, x = W.water_clusters(64, 1, seed=0)
    out.append(case("pm3_water64", sp, x, "PM3", generator=dict(fn="water_clusters", args=[64, 1], seed=0)))
    # methane and benzene through every method (Hf, which the reference's goldens do not hold)
    sp, x = read_xyz(["methane.xyz", "benzene.xyz"])
    for method in ("MNDO", "AM1", "PM3"):
        out.append(case(f"{method.lower()}_methane_benzene", sp, x, method, eps=1.0e-8, converger=(1,)))
    return out


def main():
    torch.set_default_dtype(torch.float64)  # the tables are built in the default dtype
    Force, Molecule, Constants = import_reference()
    const = Constants()
    tables = {k: getattr(const, k).tolist() for k in
              ("tore", "ussc", "uppc", "gssc", "gspc", "gppc", "gp2c", "hspc", "eheat")}
    res = []
    for c in cases():
        r = run(Force, Molecule, Constants, c.pop("_species"), c.pop("_coords"), c["method"], c["eps"],
                c["converger"])
        c.update(r)
        print(c["name"], "iterations", r["scf_iterations"], "Hf", np.round(r["Hf"], 6)[:4], flush=True)
        res.append(c)
    out = dict(source="lanl/PYSEQM seqm.basics.Force(do_force=False), float64, CPU",
               constants=tables, cases=res)
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as fh:
        json.dump(out, fh)


if __name__ == "__main__":
    main()
