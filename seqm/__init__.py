"""The lanl/PYSEQM import paths on this build: ``import seqm``, ``from seqm.Molecule import
Molecule``, ``from seqm.seqm_functions.constants import Constants`` and the other module names of
the closed-shell MNDO/AM1/PM3 path resolve to pyseqm_amd's modules, so code written against the
reference runs unchanged on the MI355X engine.  Names outside that path (excited states,
Langevin and nonadiabatic dynamics, geometry optimisation, PM6) do not exist here.
"""
import importlib
import sys

from pyseqm_amd import (Constants, Electronic_Structure, KSA_XL_BOMD, Molecular_Dynamics_Basic,  # noqa: F401
                        Molecule, XL_BOMD)
from pyseqm_amd.seqm_functions.read_xyz import read_xyz  # noqa: F401
from pyseqm_amd.xlbomd import EnergyXL, ForceXL  # noqa: F401

# reference module name -> module of this build
_MODULES = {
    "seqm.Molecule": "pyseqm_amd.Molecule",
    "seqm.ElectronicStructure": "pyseqm_amd.ElectronicStructure",
    "seqm.basics": "pyseqm_amd.basics",
    "seqm.MolecularDynamics": "pyseqm_amd.MolecularDynamics",
    "seqm.dynamics.xlbomd": "pyseqm_amd.xlbomd",
    "seqm.XLBOMD": "pyseqm_amd.xlbomd",
    "seqm.seqm_functions": "pyseqm_amd.seqm_functions",
    "seqm.seqm_functions.constants": "pyseqm_amd.seqm_functions.constants",
    "seqm.seqm_functions.parameters": "pyseqm_amd.seqm_functions.parameters",
    "seqm.seqm_functions.scf_loop": "pyseqm_amd.seqm_functions.scf_loop",
    "seqm.seqm_functions.fock": "pyseqm_amd.seqm_functions.fock",
    "seqm.seqm_functions.hcore": "pyseqm_amd.seqm_functions.hcore",
    "seqm.seqm_functions.diag": "pyseqm_amd.seqm_functions.diag",
    "seqm.seqm_functions.pack": "pyseqm_amd.seqm_functions.pack",
    "seqm.seqm_functions.SP2": "pyseqm_amd.seqm_functions.SP2",
    "seqm.seqm_functions.energy": "pyseqm_amd.seqm_functions.energy",
    "seqm.seqm_functions.two_elec_two_center_int": "pyseqm_amd.seqm_functions.two_elec_two_center_int",
    "seqm.seqm_functions.read_xyz": "pyseqm_amd.seqm_functions.read_xyz",
    "seqm.io.xyz": "pyseqm_amd.seqm_functions.read_xyz",
}
for _name, _target in _MODULES.items():
    sys.modules.setdefault(_name, importlib.import_module(_target))
sys.modules.setdefault("seqm.api", sys.modules[__name__])
sys.modules.setdefault("seqm.dynamics", sys.modules["seqm.dynamics.xlbomd"])
sys.modules.setdefault("seqm.io", sys.modules["seqm.io.xyz"])
seqm_functions = sys.modules["seqm.seqm_functions"]

__all__ = ["Constants", "Electronic_Structure", "EnergyXL", "ForceXL", "KSA_XL_BOMD", "Molecular_Dynamics_Basic",
           "Molecule", "XL_BOMD", "read_xyz"]
