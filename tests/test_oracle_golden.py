"""Oracle vs the reference's own golden vectors (tests/reference/*.json of lanl/PYSEQM).

Settings are the ones the goldens were generated with: adaptive mixing [1], eps 1e-6
(single points, tests/unit/test_smoke_single_point.py:14-16) or 1e-7 (force tests,
tests/unit/test_force_methods.py:27-30).
"""
import numpy as np
import pytest

from oracle import seqm_ref as R
from tests.conftest import golden, read_xyz


@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
def test_methane_single_point(method):
    sp, c = read_xyz(["methane.xyz"])
    r = R.single_point(sp, c, method, eps=1e-6, converger=(1,))
    g = golden(f"smoke_single_point_{method}")
    for k in ("Etot", "Eelec", "Enuc"):
        assert abs(r[k][0] - g[k]) < 1e-9, (k, r[k][0], g[k])
    f = R.forces_fd(sp, c, method, r["P"])
    np.testing.assert_allclose(f, np.array(g["force"]).reshape(f.shape), atol=1e-7, rtol=0)


def test_batch_methane_benzene_am1():
    sp, c = read_xyz(["methane.xyz", "benzene.xyz"])
    r = R.single_point(sp, c, "AM1", eps=1e-6, converger=(1,))
    g = golden("batch_single_point_am1")
    np.testing.assert_allclose(r["Etot"], g["Etot"], atol=1e-9, rtol=0)
    f = R.forces_fd(sp, c, "AM1", r["P"])
    np.testing.assert_allclose(f, np.array(g["force"]).reshape(f.shape), atol=1e-7, rtol=0)


@pytest.mark.parametrize("files,gname", [
    (["methane.xyz"], "ground_force_methane"),
    (["methanal.1.xyz", "methanal.2.xyz", "methanal.3.xyz"], "ground_force_batch_methanal"),
    (["methane.xyz", "benzene.xyz"], "ground_force_batch_mixed"),
])
def test_ground_forces_am1(files, gname):
    sp, c = read_xyz(files)
    r = R.single_point(sp, c, "AM1", eps=1e-7, converger=(1,))
    f = R.forces_fd(sp, c, "AM1", r["P"])
    np.testing.assert_allclose(f, np.array(golden(gname)["force"]).reshape(f.shape), atol=1e-7, rtol=0)


def test_odd_electron_error_message():
    sp, c = read_xyz(["methane.xyz"])
    with pytest.raises(ValueError) as ei:
        R.parse(sp, c, "AM1", charges=[1])
    assert str(ei.value) == golden("rhf_odd_electron_error")["message"]


def test_species_must_be_sorted():
    sp, c = read_xyz(["methane.xyz"])
    with pytest.raises(ValueError, match="non-increasing"):
        R.parse(sp[:, ::-1].copy(), c, "AM1")
