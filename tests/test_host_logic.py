"""Host-side logic that needs no GPU: eigensolve size groups, XL-BOMD coefficient regrouping,
serving footprints and the request stream."""
import numpy as np
import pytest

from oracle import seqm_ref as R
from pyseqm_amd.engine import size_groups
from pyseqm_amd.xlbomd import xl_coefficients


def test_size_groups_partition_and_bound_padding():
    rng = np.random.default_rng(0)
    n = rng.integers(10, 400, size=777)
    g = size_groups(n, ratio=0.8, min_group=32)
    allidx = np.sort(np.concatenate(g))
    np.testing.assert_array_equal(allidx, np.arange(len(n)))
    for k, sel in enumerate(g):
        assert len(sel) >= 32 or len(g) == 1
    # padded work never exceeds one group per molecule at the batch maximum
    work = sum(len(s) * n[s].max() ** 3 for s in g)
    assert work <= len(n) * n.max() ** 3
    assert size_groups(np.array([5])) [0].tolist() == [0]


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7, 8, 9])
def test_xl_regrouped_coefficients_equal_raw_recursion(k):
    """The regrouped coefficients used on the device (coeff_D, coeff[cindx:cindx+m] against the
    ring) reproduce the oracle's raw recursion 2P(n) - P(n-1) + kappa (D' - P(n)) + alpha sum c_i
    P(n-i) at every ring position."""
    rng = np.random.default_rng(k)
    cD, m, coeff = xl_coefficients(k)
    Pt = rng.normal(size=(m, 6))
    P = Pt[0].copy()
    Pt_o = Pt.copy()
    for step in range(2 * m):
        D = rng.normal(size=6)
        cindx = step % m
        new = cD * (0.95 * D + 0.05 * P) + (coeff[cindx:cindx + m, None] * Pt).sum(0)
        Pt[m - 1 - cindx] = new
        Po, Pt_o = R.xl_propagate(D, P, Pt_o, k, step)
        np.testing.assert_allclose(new, Po, rtol=0, atol=1e-12 * max(1.0, np.abs(Po).max()))
        P = new


def test_serving_stream_and_footprint():
    from pyseqm_amd import workloads as W
    from pyseqm_amd.serving import Request, footprint_bytes
    s = W.serving_stream(rate=1000.0, window=0.05, seed=1, water_max=12, pool=4)
    assert len(s) > 20 and all(s[i][0] <= s[i + 1][0] for i in range(len(s) - 1))
    methods = {m for _, _, _, m in s}
    assert methods <= {"AM1", "PM3"}
    for ta, sp, x, m in s:
        r = Request(0, sp, x, m)
        assert np.all(r.species[:-1] >= r.species[1:])  # the reference's atom order
        assert r.norb == 4 * (r.species > 1).sum() + (r.species == 1).sum()
    assert footprint_bytes(1536, 768 * 767 // 2, "sp2") > footprint_bytes(100, 1000, "sp2") > 0
