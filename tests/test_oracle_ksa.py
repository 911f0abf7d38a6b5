"""The oracle's thermal density and KSA-XL-BOMD pieces, pinned without the reference's code.

* fermi_q (fermi_q.py:8-72): at T -> 0 the Fermi density is the eigh projector; trace 2 nocc,
  [F, D] = 0, sum f = nocc to the reference's 1e-9; the chemical potential equals a bisection
  root of the occupation sum; the entropy of a block with two fractional levels equals the
  closed form -kB sum [f ln f + (1-f) ln(1-f)].  (Inside a wide gap the occupation sum is flat
  and mu is not determined: it is compared where sum f (1 - f) / kT is not negligible.)
* canon_dm_prt (canon_dm_prt.py:6-39): the first-order response of the trace-conserving Fermi
  density, against a central finite difference of the density itself (at 1500 K with the
  reference's 10 recursion steps, and at 30000 K, where occupations are fractional, with 20).
  Dropping the chemical-potential shift or transposing the recursion's row / column factors
  fails it.
* ksa_rank_kernel (xlbomd.py:243-339): the response W_1 = J V_1 - V_1, J the derivative of
  P -> rho(F(P)), rho = D / 2 (finite differences of fermi_q over fock), and at full rank the
  returned dP2dt2 solves (J - I)(-dP2dt2) = D - P to the fit's own residual.
"""
import numpy as np
import pytest

from oracle import seqm_ref as R
from tests.conftest import read_xyz


def _setup(names=("methane.xyz", "benzene.xyz"), method="AM1"):
    sp, c = read_xyz(list(names))
    r = R.single_point(sp, c, method, eps=1e-8, converger=(1,))
    b = r["batch"]
    H, Tt, _, _, _ = R.hcore(b)
    return b, H, Tt, r["P"]


def _mu_bisect(w, N, beta):
    lo, hi = w.min() - 100.0, w.max() + 100.0
    for _ in range(200):
        mu = 0.5 * (lo + hi)
        if R._sigmoid(-beta * (w - mu)).sum() > N:
            hi = mu
        else:
            lo = mu
    return mu


def _fermi_block(Fb, N, T):
    w, V = np.linalg.eigh(Fb)
    mu = _mu_bisect(w, N, 1.0 / (R.KB_EV * T))
    return (V * R._sigmoid(-(w - mu) / (R.KB_EV * T))) @ V.T


def test_fermi_density_limits_and_invariants():
    b, H, Tt, P = _setup()
    F = R.fock(b, P, H, Tt)
    _, Pe, _ = R.density_eigh(b, F)
    D1, S1, *_ = R.fermi_q(b, F, 1.0)
    assert np.abs(D1 - Pe).max() < 1e-12 and np.abs(S1).max() == 0.0
    for T in (1500.0, 30000.0):
        D, S, es, Qs, fs, mus = R.fermi_q(b, F, T)
        for m in range(b.nmol):
            assert abs(fs[m].sum() - b.nocc[m]) <= 1e-9
            assert abs(np.trace(D[m]) - 2 * b.nocc[m]) < 1e-8
            assert np.abs(D[m] @ F[m] - F[m] @ D[m]).max() < 1e-9
            beta = 1.0 / (R.KB_EV * T)
            if (beta * fs[m] * (1 - fs[m])).sum() > 1e-3:  # mu is pinned down (inside a wide gap it is not)
                assert abs(mus[m] - _mu_bisect(es[m], b.nocc[m], beta)) < 1e-8
        assert (S[1] > 0.0) if T > 10000 else True


def test_entropy_closed_form_on_two_fractional_levels():
    """A 4-orbital block with HOMO and LUMO 0.02 eV apart at 300 K: two fractional occupations."""
    class B:  # the fields fermi_q reads
        nmol = 1
        nocc = np.array([2])
        atom_mol = np.array([0])
        atom_loc = np.array([0])
        Z = np.array([6])
    w = np.array([-10.0, -0.01, 0.01, 5.0])
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)))
    F = (Q * w) @ Q.T
    T = 300.0
    D, S, es, Qs, fs, mus = R.fermi_q(B, F[None], T)
    f = R._sigmoid(-(w - mus[0]) / (R.KB_EV * T))
    ok = (f > 1e-14) & (1 - f > 1e-14)
    fo = f[ok]
    Sref = -R.KB_EV * np.sum(fo * np.log(fo) + (1 - fo) * np.log(1 - fo))
    assert abs(S[0] - Sref) < 1e-15 and S[0] > 1e-5
    np.testing.assert_allclose(D[0], 2.0 * (Q * f) @ Q.T, atol=1e-13)


@pytest.mark.parametrize("T,m,tol", [(1500.0, 10, 1e-9), (30000.0, 20, 1e-9)])
def test_canon_dm_prt_is_the_fermi_density_response(T, m, tol):
    b, H, Tt, P = _setup()
    F = R.fock(b, P, H, Tt)
    D, S, es, Qs, fs, mus = R.fermi_q(b, F, T)
    rng = np.random.default_rng(1)
    for mo in range(b.nmol):
        idx = R.orbital_index(b, mo)
        ix = np.ix_(idx, idx)
        X = rng.normal(size=(len(idx), len(idx)))
        X = X + X.T
        h = 1e-5
        fd = (_fermi_block(F[mo][ix] + h * X, b.nocc[mo], T) - _fermi_block(F[mo][ix] - h * X, b.nocc[mo], T)) / (2 * h)
        p1 = R.canon_dm_prt(X, T, Qs[mo], es[mo], mus[mo], m=m)
        assert np.abs(p1 - fd).max() < tol, (mo, np.abs(p1 - fd).max())
        if T > 10000:  # occupations fractional: the chemical-potential shift matters
            beta = 1.0 / (R.KB_EV * T)
            cn = 2.0 ** (-2 - m) * beta
            Xe = Qs[mo].T @ X @ Qs[mo]
            p0 = 0.5 - cn * (es[mo] - mus[mo])
            P1 = -cn * Xe
            for _ in range(m):
                p02 = p0 * p0
                dX1 = p0[:, None] * P1 + P1 * p0[None, :]
                iD0 = 1.0 / (2.0 * (p02 - p0) + 1.0)
                p0 = iD0 * p02
                P1 = iD0[:, None] * (dX1 + 2.0 * (P1 - dX1) * p0[None, :])
            noshift = Qs[mo] @ P1 @ Qs[mo].T
            assert np.abs(noshift - fd).max() > 1e-4


def test_krylov_responses_and_full_rank_solution():
    """Methane alone (8 orbitals).  The kernel's J is the reference's: V -> the occupation-scale
    response rho'(F) G(V) of rho = D / 2 to the Fock change of a density-scale V (EnergyXL applies
    canon_dm_prt to G(V) without a factor 2).  By finite differences of P -> rho(F(P)):
    W_1 = J V_1 - V_1, and with rank 8 the fit is exact and -dP2dt2 solves (J - I) x = D - P."""
    b, H, Tt, P0 = _setup(("methane.xyz",))
    T = 1500.0
    rng = np.random.default_rng(7)
    idx = R.orbital_index(b, 0)
    ix = np.ix_(idx, idx)
    A = np.zeros_like(P0)
    A[0][ix] = rng.normal(size=(len(idx), len(idx)))
    P = P0 + 0.02 * (A + A.transpose(0, 2, 1))  # away from self-consistency: D - P is not small

    def rho(Pp):
        D, *_ = R.fermi_q(b, R.fock(b, Pp, H, Tt), T)
        return D / 2.0

    F = R.fock(b, P, H, Tt)
    D, S, es, Qs, fs, mus = R.fermi_q(b, F, T)
    dP2dt2, err, rank = R.ksa_rank_kernel(b, D, P, Qs, es, mus, T, Tt, 8, 0.0)
    h = 1e-5
    V = (D - P) / np.sqrt(np.sum((D - P) ** 2))  # the first Krylov vector
    JV = (rho(P + h * V) - rho(P - h * V)) / (2 * h)
    W = R.canon_dm_prt(R.g_matrix(b, V, Tt)[0][ix], T, Qs[0], es[0], mus[0])
    assert np.abs(W - JV[0][ix]).max() < 1e-8
    assert rank == 8 and err.max() < 1e-8
    x = -dP2dt2
    Jx = (rho(P + h * x) - rho(P - h * x)) / (2 * h)
    np.testing.assert_allclose(Jx - x, D - P, atol=1e-7)
