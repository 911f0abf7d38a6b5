"""The reference's import paths resolve to this build (CPU: no compute calls)."""
import numpy as np

import pyseqm_amd


def test_reference_module_names_resolve():
    import seqm
    from seqm.ElectronicStructure import Electronic_Structure
    from seqm.Molecule import Molecule
    from seqm.MolecularDynamics import XL_BOMD
    from seqm.seqm_functions.constants import Constants
    from seqm.seqm_functions.read_xyz import read_xyz
    from seqm.seqm_functions.scf_loop import scf_loop
    from seqm.dynamics.xlbomd import ForceXL
    assert Molecule is pyseqm_amd.Molecule and Electronic_Structure is pyseqm_amd.Electronic_Structure
    assert Constants is pyseqm_amd.Constants and XL_BOMD is pyseqm_amd.XL_BOMD
    assert seqm.Molecule is Molecule and callable(scf_loop) and ForceXL is not None
    assert read_xyz is seqm.read_xyz


def test_read_xyz_sorts_and_pads():
    from seqm.seqm_functions.read_xyz import read_xyz
    from tests.conftest import GOLDEN, read_xyz as ref_style
    files = ["methanal.1.xyz", "methane.xyz", "benzene.xyz"]
    sp, x = read_xyz([f"{GOLDEN}/data/{f}" for f in files])
    sp2, x2 = ref_style(files)
    np.testing.assert_array_equal(sp, sp2)
    np.testing.assert_array_equal(x, x2)
    assert all(np.all(np.diff(row[row > 0]) <= 0) for row in sp)
