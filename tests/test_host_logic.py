"""Host-side logic that needs no GPU: eigensolve size groups, XL-BOMD coefficient regrouping,
serving footprints, the request stream and the server's admission / batching."""
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


def test_size_groups_defaults_on_the_alkane_batch():
    """The default grouping on the 1024-alkane batch (14-182 orbitals): every molecule once, at least
    two groups, each group spans down to 0.7 of its largest (sizes below the solver floor count as the
    floor), and the padded O(n^3) work is below padding every block to the largest."""
    from pyseqm_amd import workloads as W
    from pyseqm_amd.engine import EIGH_MIN_N
    sp, _ = W.alkanes(1024, seed=0)
    norb = 4 * (sp > 1).sum(1) + (sp == 1).sum(1)
    g = size_groups(norb)
    np.testing.assert_array_equal(np.sort(np.concatenate(g)), np.arange(len(norb)))
    assert len(g) >= 2
    eff = np.maximum(norb, EIGH_MIN_N)
    for sel in g[:-1]:
        assert eff[sel].min() >= 0.7 * eff[sel].max()
    work = sum(len(s) * eff[s].max() ** 3 for s in g)
    assert work < 0.6 * len(norb) * norb.max() ** 3


def test_finite_pair_cutoff_that_drops_pairs_is_refused():
    import torch
    import pyseqm_amd
    from pyseqm_amd.seqm_functions.scf_loop import scf_loop
    from tests.conftest import read_xyz
    sp, c = read_xyz(["benzene.xyz"])
    p = {"method": "AM1", "scf_eps": 1.0e-6, "scf_converger": [2], "pair_outer_cutoff": 3.0}
    mol = pyseqm_amd.Molecule(pyseqm_amd.Constants(), p, torch.as_tensor(c), torch.as_tensor(sp))
    assert len(mol.idxi) < 66  # the parser drops the pairs beyond 3 A, as the reference's does
    with pytest.raises(NotImplementedError, match="pair_outer_cutoff"):
        scf_loop(mol, eps=1e-6)


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


class _Req:
    def __init__(self, rid, method, norb, nbytes):
        self.rid, self.method, self.norb, self.bytes = rid, method, norb, nbytes


class _Srv:
    """The attributes SCFServer._take reads, without a device."""

    def __init__(self, queue, budget=10**9, batch_bytes=10**8, max_batch=8):
        import collections
        self.queue = collections.deque(queue)
        self.hbm_budget, self.batch_bytes, self.max_batch = budget, batch_bytes, max_batch
        self.in_use, self.peak_in_use = 0, 0
        self.size_class = SCFServer.size_class


from pyseqm_amd.serving import SCFServer  # noqa: E402


def test_serving_batches_keep_method_and_size_class_in_arrival_order():
    q = [_Req(0, "PM3", 300, 10), _Req(1, "AM1", 50, 1), _Req(2, "PM3", 100, 2), _Req(3, "PM3", 900, 30),
         _Req(4, "PM3", 260, 9), _Req(5, "AM1", 80, 1), _Req(6, "PM3", 40, 1)]
    srv = _Srv(q)
    batch, total = SCFServer._take(srv)
    # the oldest request leads; same method and class (> 256 orbitals) follow in arrival order
    assert [r.rid for r in batch] == [0, 3, 4] and total == 49
    assert [r.rid for r in srv.queue] == [1, 2, 5, 6]
    assert srv.in_use == 49
    batch, _ = SCFServer._take(srv)
    assert [r.rid for r in batch] == [1, 5]
    batch, _ = SCFServer._take(srv)
    assert [r.rid for r in batch] == [2, 6] and not srv.queue


def test_serving_admission_respects_budgets():
    q = [_Req(k, "PM3", 100, 40) for k in range(5)]
    srv = _Srv(q, budget=100, batch_bytes=90)
    batch, total = SCFServer._take(srv)
    assert [r.rid for r in batch] == [0, 1] and total == 80      # batch cap 90
    assert SCFServer._take(srv) is None                          # 80 in flight + 40 > budget 100
    srv.in_use = 0
    batch, _ = SCFServer._take(srv)
    assert [r.rid for r in batch] == [2, 3]


def test_sp2_launch_order_is_balanced_largest_first():
    """The SP2 launch order: a permutation of the active molecules, dealt into 8 runs (one per XCD) of
    sizes differing by at most one, each run in non-increasing orbital count, and run x holding the
    (x + 8 r)-th largest molecules (so every run gets an even share of the work)."""
    import types
    from pyseqm_amd.sp2 import _xcd_lpt_order
    rng = np.random.default_rng(3)
    norb = rng.integers(14, 183, size=300)
    for nact in (9, 64, 203, 300):
        act = np.sort(rng.choice(300, size=nact, replace=False))
        eng = types.SimpleNamespace(_idx_host=act, b=types.SimpleNamespace(norb=norb), device="cpu")
        perm = _xcd_lpt_order(eng).numpy()
        assert perm.dtype == np.int32 and sorted(perm.tolist()) == act.tolist()
        srt = act[np.argsort(-norb[act], kind="stable")]
        runs, pos = [], 0
        for x in range(8):
            r = perm[pos:pos + len(srt[x::8])]
            pos += len(r)
            assert r.tolist() == srt[x::8].tolist()
            assert np.all(np.diff(norb[r]) <= 0)
            runs.append(len(r))
        assert max(runs) - min(runs) <= 1


def test_request_records_round_trip():
    from pyseqm_amd import wire
    from pyseqm_amd import workloads as W
    sp, x = W.alkanes(3, seed=1, nmax=5)
    ws, wx = W.water_clusters(4, 1, seed=2)
    rng = np.random.default_rng(0)
    recs = []
    buf = b""
    for m in range(3):
        n = int((sp[m] > 0).sum())
        perm = rng.permutation(n)  # any atom order on the wire
        buf += wire.encode(sp[m, :n][perm], x[m, :n][perm], "AM1", 0)
    buf += wire.encode(ws[0], wx[0], "PM3", -2)
    recs = wire.decode(buf)
    assert [r[2] for r in recs] == ["AM1"] * 3 + ["PM3"] and recs[3][3] == -2
    for m in range(3):
        n = int((sp[m] > 0).sum())
        np.testing.assert_array_equal(recs[m][0], sp[m, :n])  # decreasing Z, stable
        assert {tuple(v) for v in recs[m][1]} == {tuple(v) for v in x[m, :n]}
    s2, x2, ch = wire.batch(recs[:3])
    assert s2.shape[0] == 3 and ch.tolist() == [0, 0, 0]
    with pytest.raises(ValueError, match="truncated"):
        wire.decode(buf[:-5])
    with pytest.raises(ValueError, match="request record"):
        wire.decode(b"XXXX" + buf[4:])


def test_trained_parameters_selects_per_atom_tensors_that_ask_for_gradients():
    """Energy / scf_loop wrap their outputs for autograd exactly when a learned per-atom
    parameter requires grad (and grad mode is on); other per-atom tensors and plain values pass
    through, and names the device gradient does not cover are refused rather than dropped."""
    import torch

    from pyseqm_amd.basics import _trained_parameters
    a = torch.ones(3, dtype=torch.float64, requires_grad=True)
    b = torch.ones(3, dtype=torch.float64)
    assert _trained_parameters({}) == []
    assert _trained_parameters({"g_ss": b, "U_ss": 1.0}) == []
    got = _trained_parameters({"g_ss": a, "beta_s": b})
    assert [n for n, _ in got] == ["g_ss"] and got[0][1] is a
    with torch.no_grad():
        assert _trained_parameters({"g_ss": a}) == []
    with pytest.raises(NotImplementedError, match="Kbeta"):
        _trained_parameters({"Kbeta": torch.ones(2, 2, requires_grad=True)})


def _sturm_ratio(d, e2, x, pivmin):
    q = d[0] - x
    q = -pivmin if abs(q) < pivmin else q
    cnt = int(q < 0.0)
    for k in range(1, len(d)):
        q = (d[k] - x) - e2[k - 1] / q
        q = -pivmin if abs(q) < pivmin else q
        cnt += int(q < 0.0)
    return cnt


def _sturm_sequence(d, e2, x, pivmin):
    """The three-term form the workgroup eigensolver's multisection uses (eigh.hip sturm_count)."""
    import math
    p0, p1 = 1.0, d[0] - x
    p1 = -pivmin if abs(p1) < pivmin else p1
    cnt = int(p1 < 0.0)
    for k in range(1, len(d)):
        p2 = (d[k] - x) * p1 - e2[k - 1] * p0
        if abs(p2) < pivmin * abs(p1):
            p2 = -pivmin * p1
        cnt += int(math.copysign(1.0, p2) != math.copysign(1.0, p1))
        p0, p1 = p1, p2
        if k % 8 == 0:
            ex = math.frexp(max(abs(p0), abs(p1)))[1]
            p0, p1 = math.ldexp(p0, -ex), math.ldexp(p1, -ex)
    return cnt


def test_sturm_sequence_counts_equal_the_ratio_form_and_the_spectrum():
    """Sign changes of p_k = (d_k - x) p_{k-1} - e_{k-1}^2 p_{k-2} (rescaled by powers of two) count
    the negative LDL^T pivots of T - x I, i.e. the eigenvalues below x: decoupled blocks, equal
    diagonals, shifts on diagonal elements and tiny couplings included."""
    rng = np.random.default_rng(5)
    for trial in range(60):
        n = int(rng.integers(1, 80))
        d = rng.uniform(-50, 50, n)
        e = rng.uniform(-10, 10, n)
        mode = trial % 4
        if mode == 1:
            e[rng.random(n) < 0.25] = 0.0
        elif mode == 2:
            d[:] = 3.0
        elif mode == 3:
            e[rng.random(n) < 0.3] *= 1e-9
        e2 = e[:-1] ** 2
        pivmin = 1e-300 * max(1.0, float(e2.max()) if n > 1 else 1.0)
        T = np.diag(d) + np.diag(e[:-1], 1) + np.diag(e[:-1], -1)
        lam = np.linalg.eigvalsh(T)
        for x in list(rng.uniform(-80, 80, 6)) + [float(d[0]), 3.0]:
            a = _sturm_ratio(d, e2, x, pivmin)
            b = _sturm_sequence(d, e2, x, pivmin)
            assert a == b, (trial, x, a, b)
            if np.min(np.abs(lam - x)) > 1e-9 * max(1.0, np.abs(lam).max()):
                assert b == int((lam < x).sum()), (trial, x, b)
