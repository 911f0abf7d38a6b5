"""The oracle's heat of formation and absolute energies against vectors written by the reference.

tests/golden/reference_vectors.json comes from tests/golden/make_reference_vectors.py, which runs
lanl/PYSEQM itself on the CPU (the energy module of Force(do_force=False), seqm/basics.py:1278)
and stores Hf, Etot, Eelec, Enuc, Eiso per molecule and the constants tables behind Hf
(seqm/seqm_functions/constants.py, energy.py:8-23, 195-216).

Independently of the reference, the isolated-atom coefficients (ussc .. hspc) are derived here
from the ground configuration of each element: s^a p^b with Hund's rule, counting Coulomb pairs
and same-spin exchange pairs over the one-center integrals.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import seqm_ref as R
from pyseqm_amd import workloads as W
from pyseqm_amd.seqm_functions import constants as C
from tests.conftest import GOLDEN

VEC = json.load(open(os.path.join(GOLDEN, "reference_vectors.json")))
CASES = {c["name"]: c for c in VEC["cases"]}


def case_input(c):
    """Species and coordinates of a case: stored whole, or regenerated and checked by checksum."""
    if "coordinates" in c:
        return np.array(c["species"], dtype=np.int64), np.array(c["coordinates"])
    g = c["generator"]
    sp, x = getattr(W, g["fn"])(*g["args"], seed=g["seed"])
    if "stop" in g:
        sp, x = sp[:g["stop"]], x[:g["stop"]]
    K = int((sp > 0).sum(1).max())
    sp, x = sp[:, :K], x[:, :K]
    np.testing.assert_allclose([x.sum(), (x * x).sum()], c["coord_checksum"], rtol=1e-13, atol=1e-9)
    return sp, x


def test_constants_tables_match_reference():
    t = VEC["constants"]
    n = len(C.TORE)
    for ours, key in ((C.TORE, "tore"), (C.USSC, "ussc"), (C.UPPC, "uppc"), (C.GSSC, "gssc"), (C.GSPC, "gspc"),
                      (C.GPPC, "gppc"), (C.GP2C, "gp2c"), (C.HSPC, "hspc")):
        np.testing.assert_array_equal(np.array(ours, float), np.array(t[key][:n]), err_msg=key)
    # the reference stores eheat in eV (kcal/mol / 23.061, constants.py:213)
    np.testing.assert_allclose(np.array(C.EHEAT) / C.ev_kcalpmol, t["eheat"][:n], rtol=0, atol=1e-15)
    np.testing.assert_allclose(R.EHEAT_KCAL / R.EV_KCAL, t["eheat"][:n], rtol=0, atol=1e-15)
    for ours, key in ((R.USSC, "ussc"), (R.UPPC, "uppc"), (R.GSSC, "gssc"), (R.GSPC, "gspc"), (R.GPPC, "gppc"),
                      (R.GP2C, "gp2c"), (R.HSPC, "hspc"), (R.TORE, "tore")):
        np.testing.assert_array_equal(ours, np.array(t[key][:n]), err_msg=key)


def ground_configuration_coefficients(nval):
    """(uss, upp, gss, gpp, gsp, gp2, hsp) coefficients of the isolated-atom energy for nval valence
    electrons in s^min(nval,2) p^rest, p shells filled by Hund's rule (one spin first over px, py, pz).

    Pairs of electrons: s-s (opposite spins) -> gss; s-p -> gsp, minus hsp if the spins agree;
    p-p in one orbital -> gpp; p-p in different orbitals -> gp2, minus the exchange
    hpp = (gpp - gp2)/2 if the spins agree."""
    ns = min(nval, 2)
    npl = nval - ns
    s_el = [("s", 0, +1), ("s", 0, -1)][:ns]
    p_el = [("p", k % 3, +1 if k < 3 else -1) for k in range(npl)]
    c = dict(uss=ns, upp=npl, gss=0.0, gpp=0.0, gsp=0.0, gp2=0.0, hsp=0.0)
    for a, b in itertools.combinations(s_el + p_el, 2):
        same_spin = a[2] == b[2]
        if a[0] == "s" and b[0] == "s":
            c["gss"] += 1
        elif a[0] != b[0]:
            c["gsp"] += 1
            c["hsp"] -= same_spin
        elif a[1] == b[1]:
            c["gpp"] += 1
        else:
            c["gp2"] += 1
            if same_spin:  # - hpp = -(gpp - gp2)/2
                c["gpp"] -= 0.5
                c["gp2"] += 0.5
    return c


@pytest.mark.parametrize("Z", [1, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16, 17])
def test_isolated_atom_coefficients_follow_ground_configuration(Z):
    c = ground_configuration_coefficients(int(C.TORE[Z]))
    got = dict(uss=C.USSC[Z], upp=C.UPPC[Z], gss=C.GSSC[Z], gpp=C.GPPC[Z], gsp=C.GSPC[Z], gp2=C.GP2C[Z],
               hsp=C.HSPC[Z])
    assert got == pytest.approx(c), (Z, got, c)


def test_eiso_per_molecule_matches_reference():
    """Eiso summed per molecule (energy.py:195-216) from the oracle's parameters and tables."""
    for c in VEC["cases"]:
        if "coordinates" not in c:
            continue
        sp, x = case_input(c)
        b = R.parse(sp, x, c["method"])
        e = np.zeros(b.nmol)
        np.add.at(e, b.atom_mol, R.eiso(b))
        np.testing.assert_allclose(e, c["Eiso"], rtol=0, atol=1e-10, err_msg=c["name"])


CASE_NAMES = ["mndo_methane_benzene", "am1_methane_benzene", "pm3_methane_benzene", "pm3_water8", "mndo_water8",
              "am1_water8", "pm3_water16", "pm3_water64", "am1_alkanes_c2_slice", "pm3_benzene16_c1",
              "am1_alkanes_c2_first128"]


@pytest.mark.parametrize("name", CASE_NAMES)
def test_oracle_matches_reference_vectors(name):
    """Hf, Etot, Eelec, Enuc of the oracle against the reference run at the same eps and mixer, at
    1e-6 eV (measured: <= 6e-10 eV), and the same number of SCF iterations (the reference's printed
    iteration counter, scf_loop.py:601, 978)."""
    c = CASES[name]
    sp, x = case_input(c)
    r = R.single_point(sp, x, c["method"], eps=c["eps"], converger=tuple(c["converger"]))
    assert not r["notconverged"].any()
    for k in ("Hf", "Etot", "Eelec", "Enuc"):
        np.testing.assert_allclose(r[k], c[k], rtol=0, atol=1e-6, err_msg=f"{name} {k}")
    assert r["iterations"] == c["scf_iterations"], (r["iterations"], c["scf_iterations"])
