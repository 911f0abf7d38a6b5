"""Pins of the oracle against what the mathematics fixes, independent of how the oracle
(and the reference) write each formula.

* local two-center integrals  vs a direct point-charge (multipole) Klopman-Ohno sum
* STO overlaps                vs numerical quadrature of normalised Slater orbitals
* rho1 / rho2                 vs the residual of their defining equations
* one-center Fock terms       vs the general NDDO G(P) formula
* frame independence, rotation/translation invariance, density-step invariants
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import seqm_ref as R
from tests.conftest import read_xyz

EV = R.EV


# ------------------------------------------------------------ point charges
def _charges(kind, D1, D2, r0, r1, r2):
    """Charge model of an orbital-product distribution in the local frame (sigma, pi, pi')."""
    s = [(1.0, (0.0, 0.0, 0.0), r0)]
    ax = {"o": 0, "p": 1, "q": 2}
    if kind == "ss":
        return s
    if kind[0] == "s":  # s-p dipole
        e = np.zeros(3)
        e[ax[kind[1]]] = D1
        return [(0.5, tuple(e), r1), (-0.5, tuple(-e), r1)]
    a, b = ax[kind[0]], ax[kind[1]]
    if a == b:  # monopole + linear quadrupole
        e = np.zeros(3)
        e[a] = 2.0 * D2
        return s + [(0.25, tuple(e), r2), (0.25, tuple(-e), r2), (-0.5, (0.0, 0.0, 0.0), r2)]
    out = []
    for sa in (1, -1):
        for sb in (1, -1):
            e = np.zeros(3)
            e[a], e[b] = sa * D2, sb * D2
            out.append((0.25 * sa * sb, tuple(e), r2))
    return out


_KINDS = {(0, 0): "ss", (0, 1): "so", (0, 2): "sp", (0, 3): "sq", (1, 1): "oo", (2, 2): "pp",
          (3, 3): "qq", (1, 2): "op", (1, 3): "oq", (2, 3): "pq"}


def _local_tensor_charges(r, A, B):
    L = np.zeros((4, 4, 4, 4))
    for (a, b), ka in _KINDS.items():
        for (c, d), kb in _KINDS.items():
            v = 0.0
            for qa, xa, ra in _charges(ka, *A):
                for qb, xb, rb in _charges(kb, *B):
                    dx = np.array(xa) - (np.array(xb) + np.array([-r, 0.0, 0.0]))
                    v += qa * qb * EV / math.sqrt(dx @ dx + (ra + rb) ** 2)
            for p, q in {(a, b), (b, a)}:
                for u, w in {(c, d), (d, c)}:
                    L[p, q, u, w] = v
    return L


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_local_integrals_match_point_charge_model(seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(1.5, 9.0)
    A = (rng.uniform(0.3, 1.2), rng.uniform(0.3, 1.2), rng.uniform(0.5, 1.5), rng.uniform(0.4, 1.3),
         rng.uniform(0.4, 1.3))
    B = (rng.uniform(0.3, 1.2), rng.uniform(0.3, 1.2), rng.uniform(0.5, 1.5), rng.uniform(0.4, 1.3),
         rng.uniform(0.4, 1.3))
    ri = R.local_ri_xx(np.array([r]), np.array([A[0]]), np.array([B[0]]), np.array([A[1]]),
                       np.array([B[1]]), np.array([A[2]]), np.array([B[2]]), np.array([A[3]]),
                       np.array([B[3]]), np.array([A[4]]), np.array([B[4]]))
    L = R.local_tensor_from_ri(ri)[0]
    Lc = _local_tensor_charges(r, A, B)
    # (pi pi'|pi pi') is not the square-quadrupole value: MNDO fixes it by invariance under
    # rotation about the bond axis, (pi pi'|pi pi') = [(pi pi|pi pi) - (pi pi|pi' pi')] / 2.
    # That invariance is what test_two_center_frame_independent checks.
    for idx in ((2, 3, 2, 3), (2, 3, 3, 2), (3, 2, 2, 3), (3, 2, 3, 2)):
        Lc[idx] = 0.5 * (Lc[2, 2, 2, 2] - Lc[2, 2, 3, 3])
    np.testing.assert_allclose(L, Lc, atol=1e-12 * EV, rtol=0)
    # heavy-H: B carries only the s monopole
    rxh = R.local_ri_xh(np.array([r]), np.array([A[0]]), np.array([A[1]]), np.array([A[2]]),
                        np.array([A[3]]), np.array([A[4]]), np.array([B[2]]))[0]
    np.testing.assert_allclose(rxh, [Lc[0, 0, 0, 0], Lc[0, 1, 0, 0], Lc[1, 1, 0, 0], Lc[2, 2, 0, 0]],
                               atol=1e-12 * EV, rtol=0)
    hh = R.local_ri_hh(np.array([r]), np.array([A[2]]), np.array([B[2]]))[0]
    assert abs(hh - Lc[0, 0, 0, 0]) < 1e-12 * EV


# ------------------------------------------------------------------ overlap
def _sto(n, l, zeta, x, y, z):
    r = math.sqrt(x * x + y * y + z * z)
    N = (2 * zeta) ** (n + 0.5) / math.sqrt(math.factorial(2 * n))
    rad = N * r ** (n - 1) * math.exp(-zeta * r)
    if l == "s":
        return rad / math.sqrt(4 * math.pi)
    c = {"x": x, "y": y, "z": z}[l] / r if r > 0 else 0.0
    return rad * math.sqrt(3 / (4 * math.pi)) * c


def _quad_overlap(na, la, za, nb, lb, zb, Rb):
    if la in "xy" or lb in "xy":  # pi-pi: phi integral of cos^2 = pi
        def f(z, rho):
            return math.pi * rho * (_sto(na, la, za, rho, 0, z) * _sto(nb, lb, zb, rho, 0, z - Rb))
    else:
        def f(z, rho):
            return 2 * math.pi * rho * _sto(na, la, za, 0.0, rho, z) * _sto(nb, lb, zb, 0.0, rho, z - Rb)
    val, _ = integrate.dblquad(f, 0.0, 25.0, -25.0, Rb + 25.0, epsabs=1e-11, epsrel=1e-10)
    return val


@pytest.mark.parametrize("zi,zj,rb", [(1, 1, 1.4), (6, 1, 2.05), (6, 6, 2.6), (8, 6, 2.3), (8, 1, 1.8),
                                      (16, 1, 2.5), (16, 6, 3.3), (16, 16, 3.8)])
def test_overlap_matches_quadrature(zi, zj, rb):
    prm = R.load_params("PM3")
    za = np.array([[prm["zeta_s"][zi], prm["zeta_p"][zi]]])
    zb = np.array([[prm["zeta_s"][zj], prm["zeta_p"][zj]]])
    S = R.overlap_local(np.array([zi]), np.array([zj]), np.array([rb]), za, zb)[:, 0]
    na, nb = R.QN[zi], R.QN[zj]
    ref_ss = _quad_overlap(na, "s", za[0, 0], nb, "s", zb[0, 0], rb)
    assert abs(S[0] - ref_ss) < 2e-7, (S[0], ref_ss)
    if zi > 1:  # <p_sigma(A) | s(B)> with B on +z: S211
        v = _quad_overlap(na, "z", za[0, 1], nb, "s", zb[0, 0], rb)
        assert abs(S[1] - v) < 2e-7, (S[1], v)
    if zj > 1:
        v = _quad_overlap(na, "s", za[0, 0], nb, "z", zb[0, 1], rb)  # = -S121
        assert abs(-S[2] - v) < 2e-7, (S[2], v)
        v = _quad_overlap(na, "z", za[0, 1], nb, "z", zb[0, 1], rb)  # = -S221
        assert abs(-S[3] - v) < 2e-7, (S[3], v)
        v = _quad_overlap(na, "x", za[0, 1], nb, "x", zb[0, 1], rb)  # = S222
        assert abs(S[4] - v) < 2e-7, (S[4], v)


def test_aux_integrals_match_quadrature():
    xs = np.array([0.05, 0.3, 0.49, 0.51, 1.7, 4.0, -0.3, -2.2])
    B = R.bint(xs)
    for i, x in enumerate(xs):
        for k in range(7):  # overlaps of up to third-row atoms use B_0 .. B_6
            v, _ = integrate.quad(lambda t: t ** k * math.exp(-x * t), -1, 1, epsabs=1e-14)
            # |x| <= 0.5 uses the reference's series, truncated after x^6 (even k) / x^5 (odd k):
            # the first dropped term bounds the error there.
            tol = 2.0 * abs(x) ** (8 if k % 2 == 0 else 7) if abs(x) <= 0.5 else 1e-9 * max(1.0, abs(v))
            assert abs(B[i, k] - v) < tol, (x, k, B[i, k], v)
    xa = np.array([0.7, 2.0, 6.5])
    A = R.aint(xa, np.full(3, 431))
    for i, x in enumerate(xa):
        for k in range(13):
            v, _ = integrate.quad(lambda t: t ** k * math.exp(-x * t), 1, np.inf, epsabs=1e-14)
            assert abs(A[i, k] - v) < 1e-10 * max(1.0, abs(v)), (x, k)


# --------------------------------------------------------------- rho1/rho2
@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
def test_rho_defining_equations(method):
    prm = R.load_params(method)
    Z = np.array([6, 7, 8, 16])
    dd, qq, rho0, r1, r2 = R.atom_multipoles(Z, prm)
    ad, aq = 0.5 / r1, 0.5 / r2
    hsp = prm["h_sp"][Z] / EV
    hpp = np.maximum(0.5 * (prm["g_pp"][Z] - prm["g_p2"][Z]), 0.1) / EV
    res1 = 0.5 * ad - 0.5 / np.sqrt(4 * dd ** 2 + 1 / ad ** 2) - hsp
    res2 = 0.25 * aq - 0.5 / np.sqrt(4 * qq ** 2 + 1 / aq ** 2) + 0.25 / np.sqrt(8 * qq ** 2 + 1 / aq ** 2) - hpp
    assert np.all(np.abs(res1) < 1e-9) and np.all(np.abs(res2) < 1e-9)
    np.testing.assert_allclose(rho0, 0.5 * EV / prm["g_ss"][Z])
    # dd and qq closed forms for n = 2: dd = 5 (4 zs zp)^2.5 / (zs+zp)^6 / sqrt 3; qq = sqrt(3/2)/zp
    zs, zp = prm["zeta_s"][6], prm["zeta_p"][6]
    assert abs(dd[0] - 5 * (4 * zs * zp) ** 2.5 / (zs + zp) ** 6 / math.sqrt(3)) < 1e-14
    assert abs(qq[0] - math.sqrt(1.5) / zp) < 1e-14


# ------------------------------------------------------------ one-center G
def test_one_center_fock_closed_forms():
    prm = R.load_params("AM1")
    Z = np.array([8])
    G1 = R.one_center_tensor(Z, prm)
    rng = np.random.default_rng(3)
    A = rng.normal(size=(4, 4))
    P = A + A.T
    G = np.einsum("cd,abcd->ab", P, G1[0]) - 0.5 * np.einsum("cd,acbd->ab", P, G1[0])
    gss, gsp, gpp, gp2, hsp = (prm[k][8] for k in ("g_ss", "g_sp", "g_pp", "g_p2", "h_sp"))
    pss, ppt = P[0, 0], P[1, 1] + P[2, 2] + P[3, 3]
    assert abs(G[0, 0] - (0.5 * pss * gss + ppt * (gsp - 0.5 * hsp))) < 1e-12
    for i in (1, 2, 3):
        exp = pss * (gsp - 0.5 * hsp) + 0.5 * P[i, i] * gpp + (ppt - P[i, i]) * (1.25 * gp2 - 0.25 * gpp)
        assert abs(G[i, i] - exp) < 1e-12
        assert abs(G[0, i] - P[0, i] * (1.5 * hsp - 0.5 * gsp)) < 1e-12
    for i, j in ((1, 2), (1, 3), (2, 3)):
        assert abs(G[i, j] - P[i, j] * (0.75 * gpp - 1.25 * gp2)) < 1e-12


# ---------------------------------------------------------- frame / invariance
def test_two_center_frame_independent():
    sp, c = read_xyz(["methanal.2.xyz"])
    b = R.parse(sp, c, "PM3")
    T, _, rxx = R.two_center_tensor(b)
    # complete the pi axes differently: rotate rows 1,2 of the frame by an arbitrary angle
    dd, qq, rho0, rho1, rho2 = R.atom_multipoles(b.Z, b.prm)
    XX = (b.ni > 1) & (b.nj > 1)
    Rm = R.frame_from_axis(-b.xij[XX])
    th = 0.7
    rot = np.array([[1, 0, 0], [0, math.cos(th), -math.sin(th)], [0, math.sin(th), math.cos(th)]])
    T2 = R.rotate_tensor(R.local_tensor_from_ri(rxx), np.einsum("ij,njk->nik", rot, Rm))
    np.testing.assert_allclose(T[XX], T2, atol=1e-12, rtol=0)


@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
def test_energy_rotation_translation_invariance(method):
    sp, c = read_xyz(["methanal.2.xyz", "benzene.xyz"])
    e0 = R.single_point(sp, c, method, eps=1e-9, converger=(2,))
    rng = np.random.default_rng(7)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b_, c_, d = q
    Rm = np.array([[a * a + b_ * b_ - c_ * c_ - d * d, 2 * (b_ * c_ - a * d), 2 * (b_ * d + a * c_)],
                   [2 * (b_ * c_ + a * d), a * a - b_ * b_ + c_ * c_ - d * d, 2 * (c_ * d - a * b_)],
                   [2 * (b_ * d - a * c_), 2 * (c_ * d + a * b_), a * a - b_ * b_ - c_ * c_ + d * d]])
    c2 = c @ Rm.T + np.array([1.3, -4.0, 2.2])
    c2[sp == 0] = 0.0
    e1 = R.single_point(sp, c2, method, eps=1e-9, converger=(2,))
    np.testing.assert_allclose(e1["Hf"], e0["Hf"], atol=1e-7, rtol=0)


# ----------------------------------------------------------- density step
def test_density_invariants_and_sp2():
    sp, c = read_xyz(["methanal.1.xyz", "benzene.xyz"])
    r = R.single_point(sp, c, "PM3", eps=1e-9, converger=(2,))
    b, F, P = r["batch"], r["F"], r["P"]
    for mo in range(2):
        idx = R.orbital_index(b, mo)
        p = P[mo][np.ix_(idx, idx)]
        f = F[mo][np.ix_(idx, idx)]
        assert abs(np.trace(p) - 2 * b.nocc[mo]) < 1e-10
        np.testing.assert_allclose(p @ p, 2 * p, atol=1e-8)   # idempotent (spin-summed)
        assert np.abs(f @ p - p @ f).max() < 50 * 1e-9 * 10    # converged: [F, P] ~ 0
        x, it = R.sp2(f[None], [b.nocc[mo]], eps=1e-7)
        np.testing.assert_allclose(x[0], p, atol=2e-6)


def test_pulay_and_adaptive_reach_same_fixed_point():
    sp, c = read_xyz(["benzene.xyz", "methanal.3.xyz"])
    ra = R.single_point(sp, c, "AM1", eps=1e-9, converger=(1,))
    rp = R.single_point(sp, c, "AM1", eps=1e-9, converger=(2,))
    rl = R.single_point(sp, c, "AM1", eps=1e-9, converger=(0, 0.5))
    np.testing.assert_allclose(ra["Etot"], rp["Etot"], atol=1e-7, rtol=0)
    np.testing.assert_allclose(rl["Etot"], rp["Etot"], atol=1e-7, rtol=0)
    assert rp["iterations"] < ra["iterations"]


# ---------------------------------------------------------------- XL-BOMD
def test_xl_energy_is_shadow_energy():
    """E_xl(D, P) = E(D) - 1/2 B(D - P, D - P) with E(X) = 1/2 sum X (H + F(X)) and
    B(X, X) = sum X G(X), G = F - H linear; D = P gives the SCF energy."""
    sp, c = read_xyz(["methanal.1.xyz", "benzene.xyz"])
    b = R.parse(sp, c, "AM1")
    H, T, _, _, _ = R.hcore(b)
    P = R.initial_density(b)
    rng = np.random.default_rng(3)
    dP = rng.normal(scale=0.05, size=P.shape)
    dP = dP + dP.transpose(0, 2, 1)
    mask = (np.abs(P).sum(axis=2, keepdims=True) > 0) & (np.abs(P).sum(axis=1, keepdims=True) > 0)
    D = P + dP * mask
    F = R.fock(b, P, H, T)
    Ex = R.elec_energy_xl(D, P, F, H)
    FD = R.fock(b, D, H, T)
    X = D - P
    GX = R.fock(b, X, H, T) - H
    np.testing.assert_allclose(Ex, R.elec_energy(D, FD, H) - 0.5 * np.einsum("nij,nij->n", X, GX), atol=1e-9)
    np.testing.assert_allclose(R.elec_energy_xl(P, P, F, H), R.elec_energy(P, F, H), atol=1e-10)


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7, 8, 9])
def test_xl_propagation_extrapolates_linear_sequences(k):
    """The XL coefficient sets satisfy sum c_i = 0 and sum i c_i = 0, so a density history that is
    linear in time, with D(n) = P(n), is extrapolated exactly, over every ring position."""
    rng = np.random.default_rng(k)
    X, Y = rng.normal(size=(2, 5, 5))
    m = k + 1
    Pt = np.zeros((m, 5, 5))
    # seed the ring as after m steps of the linear sequence P(n) = X + n Y, n = 0..m-1
    for n in range(m):
        Pt[m - 1 - (n % m)] = X + n * Y
    P = X + (m - 1) * Y
    for step in range(m, 3 * m):
        P, Pt = R.xl_propagate(P, P, Pt, k, step)
        np.testing.assert_allclose(P, X + step * Y, atol=1e-9 * (1 + step))
