"""pyseqm_amd: batched MNDO/AM1/PM3 closed-shell SCF on MI355X (gfx950) HIP kernels, with the
seqm.Molecule / Electronic_Structure / XL_BOMD API of lanl/PYSEQM."""
from .seqm_functions.constants import Constants  # noqa: F401
from .Molecule import Molecule  # noqa: F401
from .ElectronicStructure import Electronic_Structure  # noqa: F401
from .MolecularDynamics import KSA_XL_BOMD, Molecular_Dynamics_Basic, XL_BOMD  # noqa: F401

__all__ = ["Constants", "Molecule", "Electronic_Structure", "Molecular_Dynamics_Basic", "XL_BOMD", "KSA_XL_BOMD"]
