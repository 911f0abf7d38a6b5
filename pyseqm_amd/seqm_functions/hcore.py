"""hcore(molecule): core Hamiltonian blocks and two-center integrals with the reference's return
values (seqm/seqm_functions/hcore.py:9-179), computed by the pair-integral kernel of the
molecule's device engine.  riXH / ri (local-frame intermediates kept by the reference for its
analytic gradient) are not materialised and are returned as None."""
import torch

from .fock import matrix_to_blocks


def hcore(molecule, doTETCI=True):
    eng = molecule.engine()
    eng.set_coordinates(molecule.coordinates.detach())
    b = eng.b
    M = matrix_to_blocks(b.to_padded(eng.H), b.nmol, b.molsize)
    if not doTETCI:
        return M, None, None, None, None, None
    rho0 = eng.atom_par[:b.natoms, 26]
    ii, jj = molecule.idxi.long(), molecule.idxj.long()
    return M, eng.w_reference(), rho0[ii], rho0[jj], None, None
