"""The seqm_functions entry points with the reference's signatures (hcore, two_elec_two_center_int,
fock, sym_eig_trunc, SP2, elec_energy, pair_nuclear_energy) on the GPU, fed the oracle's inputs in
the reference's layouts and compared with the oracle's outputs.

Bars: integrals / Fock / energies 1e-12 relative to the matrix scale (FP64, summation order);
FP64 density step 1e-9; mixed-precision SP2 density 1e-5 (bf16x3 + one McWeeny step to FP64 accuracy)."""
import numpy as np
import pytest
import torch

import pyseqm_amd
from oracle import seqm_ref as R
from pyseqm_amd.seqm_functions.SP2 import SP2
from pyseqm_amd.seqm_functions.diag import sym_eig_trunc
from pyseqm_amd.seqm_functions.energy import elec_energy, pair_nuclear_energy
from pyseqm_amd.seqm_functions.fock import fock
from pyseqm_amd.seqm_functions.hcore import hcore
from pyseqm_amd import workloads as W
from pyseqm_amd.seqm_functions.pack import pack, unpack
from pyseqm_amd.seqm_functions.two_elec_two_center_int import two_elec_two_center_int
from tests.test_gpu_parity import _inputs

pytestmark = pytest.mark.gpu
DEV = "cuda"
T = lambda a, **k: torch.as_tensor(a, device=DEV, **k)


def _close(a, b, rtol):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    b = np.asarray(b)
    scale = max(np.abs(b).max(), 1.0)
    err = np.abs(a - b).max()
    assert err <= rtol * scale, (err, scale)


@pytest.mark.parametrize("name,method", [("sulfur", "PM3"), ("water8", "AM1"), ("methanal+benzene", "MNDO")])
def test_integrals_and_fock_entry_points(name, method):
    sp, c = _inputs(name)
    ob = R.parse(sp, c, method)
    H, Tt, di, gam, mp = R.hcore(ob)
    w_o = R.tensor_to_w(Tt)
    prm = ob.prm
    Z = T(ob.Z)
    g = {k: T(prm[k][ob.Z]) for k in ("zeta_s", "zeta_p", "g_ss", "g_pp", "g_sp", "g_p2", "h_sp")}
    # two_elec_two_center_int on the oracle's pair list
    w, e1b, e2a, r0i, r0j, _, _ = two_elec_two_center_int(
        None, T(ob.idxi), T(ob.idxj), T(ob.ni), T(ob.nj), T(ob.xij), T(ob.rij), Z, g["zeta_s"], g["zeta_p"], None,
        None, None, None, g["g_ss"], g["g_pp"], g["g_p2"], g["h_sp"], None, None, None, None, None, method)
    _close(w, w_o, 1e-12)
    tore = R.TORE
    for k, (mu, nu) in enumerate(zip((0, 0, 1, 0, 1, 2, 0, 1, 2, 3), (0, 1, 1, 2, 2, 2, 3, 3, 3, 3))):
        _close(e1b[:, mu, nu], -tore[ob.nj] * w_o[:, k, 0], 1e-12)
        _close(e2a[:, mu, nu], -tore[ob.ni] * w_o[:, 0, k], 1e-12)
    # hcore through a Molecule
    p = {"method": method, "scf_eps": 1e-8, "scf_converger": [2]}
    mol = pyseqm_amd.Molecule(pyseqm_amd.Constants().to(DEV), p, T(c), T(sp))
    M, w2, rho0xi, rho0xj, _, _ = hcore(mol)
    Hp = M.reshape(ob.nmol, ob.molsize, ob.molsize, 4, 4).transpose(2, 3).reshape(H.shape)
    _close(Hp, np.triu(H), 1e-12)
    _close(w2, w_o, 1e-12)
    # fock with the reference's arguments
    rng = np.random.default_rng(1)
    Pr = rng.normal(size=H.shape)
    Pr = Pr + Pr.transpose(0, 2, 1)
    for a in np.nonzero(ob.Z == 1)[0]:
        o = 4 * ob.atom_loc[a]
        Pr[ob.atom_mol[a], o + 1:o + 4, :] = 0.0
        Pr[ob.atom_mol[a], :, o + 1:o + 4] = 0.0
    F = fock(ob.nmol, ob.molsize, T(Pr), M, T(ob.maskd), T(ob.mask), T(ob.idxi), T(ob.idxj), w2, None, g["g_ss"],
             g["g_pp"], g["g_sp"], g["g_p2"], g["h_sp"], method, g["zeta_s"], g["zeta_p"], None, Z, None, None)
    Fo = R.fock(ob, Pr, H, Tt)
    _close(F, Fo, 1e-12)
    _close(elec_energy(T(Pr), F, T(np.triu(H))), R.elec_energy(Pr, Fo, H), 1e-12)
    # core-core repulsion
    alpha = T(prm["alpha"][ob.Z])
    ng = {"MNDO": 0, "AM1": 4, "PM3": 2}[method]
    par = (alpha,) if ng == 0 else (alpha,) + tuple(
        torch.stack([T(prm[f"Gaussian{k}_{x}"][ob.Z]) for k in range(1, ng + 1)], 1) for x in "KLM")
    en = pair_nuclear_energy(Z, None, ob.nmol, T(ob.ni), T(ob.nj), T(ob.idxi), T(ob.idxj), T(ob.rij), rho0xi, rho0xj,
                             None, None, w2[:, 0, 0], method=method, parameters=par)
    _close(en, R.pair_nuclear_energy(ob, gam), 1e-12)


def test_density_entry_points():
    sp, c = _inputs("methanal+benzene")
    r = R.single_point(sp, c, "PM3", eps=1e-9, converger=(2,))
    ob, F = r["batch"], r["F"]
    e_o, P_o, _ = R.density_eigh(ob, F)
    e, P, v = sym_eig_trunc(T(F), T(ob.nHeavy), T(ob.nHydro), T(ob.nocc))
    _close(P, P_o, 1e-9)
    _close(e, e_o, 1e-9)
    # pack -> SP2 -> unpack (the reference's SP2 call pattern, scf_loop.py:66-78) on same-size systems
    sp6, c6 = _inputs("benzene4")
    r6 = R.single_point(sp6, c6, "PM3", eps=1e-9, converger=(2,))
    b6, F6 = r6["batch"], r6["F"]
    nh, nH = T(b6.nHeavy), T(b6.nHydro)
    Pk = SP2(pack(T(F6), nh, nH), T(b6.nocc), 1e-4)
    Ps = unpack(Pk, nh, nH, F6.shape[-1]).cpu().numpy()
    Po = r6["P"]
    # the mixed-precision density differs from the projector at first order in its elements but
    # only at second order in the energy; idempotency and trace are restored by the FP64 step
    assert np.abs(Ps - Po).max() < 2e-4
    H6 = R.hcore(b6)[0]
    dE = np.abs(R.elec_energy(Ps, R.fock(b6, Ps, H6, R.hcore(b6)[1]), H6) - R.elec_energy(Po, F6, H6)) / 12
    assert dE.max() < 1e-6, dE
    for m in range(4):
        idx = R.orbital_index(b6, m)
        p = Ps[m][np.ix_(idx, idx)]
        assert np.abs(p @ p - 2 * p).max() < 1e-6
        assert abs(np.trace(p) - 2 * b6.nocc[m]) < 1e-8


def test_custom_ops_match_engine():
    import pyseqm_amd.ops  # noqa: F401  (registers torch.ops.pyseqm)
    from pyseqm_amd import _native as nat
    from pyseqm_amd.engine import SCFEngine
    from pyseqm_amd.ops import layout_handle
    sp, c = _inputs("sulfur")
    eng = SCFEngine(sp, c, "PM3", DEV)
    r = eng.scf(eps=1e-8)
    b = eng.b
    h = layout_handle(b)
    F1 = torch.empty_like(r.P)
    e1 = torch.zeros(b.nmol, dtype=torch.float64, device=DEV)
    eng.fock(r.P, F1, eelec=e1)
    F2 = torch.empty_like(r.P)
    e2 = torch.zeros_like(e1)
    torch.ops.pyseqm.fock_build(h, eng.atom_par, eng.w_xx, eng.w_xh, eng.w_hh, r.P, eng.H, F2, eng.part,
                                eng.tile_v, eng.atom_v, e2)
    assert torch.equal(F1, F2) and torch.equal(e1, e2)
    f1 = eng.forces(r.P)
    f2 = torch.zeros(b.natoms, 3, dtype=torch.float64, device=DEV)
    torch.ops.pyseqm.pair_forces(h, nat.METHODS["PM3"], eng.coords, eng.atom_par, r.P, 1e-3, eng.part, f2)
    assert torch.equal(f1, f2)


def test_density_ops_match_engine():
    """torch.ops.pyseqm.density_eigh_fp64 / density_sp2 give the engine's densities bit for bit."""
    import pyseqm_amd.ops  # noqa: F401
    from pyseqm_amd import sp2 as S
    from pyseqm_amd.engine import SCFEngine
    from pyseqm_amd.ops import layout_handle
    sp, c = W.water_clusters(20, 3, seed=4)
    eng = SCFEngine(sp, c, "PM3", DEV)
    r = eng.scf(eps=1e-6, converger=(2,), density="eigh")
    b = eng.b
    F = eng.energies(r.P)["F"]
    idx = torch.arange(b.nmol, dtype=torch.int32, device=DEV)
    eng._idx_host = np.arange(b.nmol)
    nmax = int(b.nmax)
    h = layout_handle(b)
    P1 = torch.zeros_like(F)
    e1, V1, _ = eng.density_eigh(F, idx, b.nmol, P1, want_evecs=True)
    P2 = torch.zeros_like(F)
    e2, V2 = torch.ops.pyseqm.density_eigh_fp64(h, F, idx, e1.shape[1], P2, eng.red)
    assert torch.equal(P1, P2) and torch.equal(e1, e2)
    P3 = torch.zeros_like(F)
    S.density_sp2(eng, F, idx, b.nmol, P3, eps=1e-4, fixed_iters=40)
    ws = eng._sp2ws
    P4 = torch.zeros_like(F)
    torch.ops.pyseqm.density_sp2(h, F, idx, nmax, ws.xoff, ws.xld, ws.Xa, ws.Xb, ws.state, ws.part, ws.Y2, eng.red,
                                 ws.ndone, P4, 1e-4, 40)
    assert torch.equal(P3, P4)


def test_diis_step_op_extrapolates():
    """torch.ops.pyseqm.diis_step: after three updates F is sum_j c_j F_j with c from the bordered system
    [B 1; 1 0][c; l] = [0; 1], B_ij = <C_i, C_j> / 2 over the commutators C_j = F_j P_j - P_j F_j (the
    reference's Pulay extrapolation, scf_loop.py:990-1035), here solved by torch."""
    import pyseqm_amd.ops  # noqa: F401
    from pyseqm_amd.engine import SCFEngine
    from pyseqm_amd.ops import layout_handle
    sp, c = W.water_clusters(8, 2, seed=6)
    eng = SCFEngine(sp, c, "PM3", DEV)
    b = eng.b
    h = layout_handle(b)
    g = torch.Generator(device=DEV).manual_seed(11)
    nf, total = 10, b.total
    Fring = torch.zeros(nf, total, dtype=torch.float64, device=DEV)
    Cring = torch.zeros_like(Fring)
    X = torch.zeros(total, dtype=torch.float64, device=DEV)
    err = torch.zeros(b.nmol, dtype=torch.float64, device=DEV)
    # the bordered matrix as the engine keeps it: lower triangle, border -1, corner 0
    emat0 = torch.tril(torch.eye(nf + 1, dtype=torch.float64, device=DEV) - 1.0)
    Bmat = emat0.expand(b.nmol, -1, -1).contiguous()
    coef = torch.zeros(b.nmol, nf, dtype=torch.float64, device=DEV)
    reset = torch.zeros(b.nmol, dtype=torch.uint8, device=DEV)
    idx = torch.arange(b.nmol, dtype=torch.int32, device=DEV)
    nmax = int(b.nmax)
    Fs, Ps = [], []
    for j in range(3):
        F = torch.zeros(total, dtype=torch.float64, device=DEV)
        P = torch.zeros_like(F)
        for m in range(b.nmol):
            n, o = int(b.norb[m]), int(b.mol_mat_off[m])
            a = torch.randn(n, n, generator=g, device=DEV, dtype=torch.float64)
            q = torch.randn(n, n, generator=g, device=DEV, dtype=torch.float64)
            F[o:o + n * n] = (a + a.T).reshape(-1)
            P[o:o + n * n] = (q + q.T).reshape(-1)
        Fs.append(F.clone())
        Ps.append(P)
        torch.ops.pyseqm.diis_step(h, idx, nmax, F, P, Fring, Cring, X, err, Bmat, coef, reset, eng.red, j, j + 1,
                                   False)
    torch.cuda.synchronize()
    assert not bool(reset.any())
    for m in range(b.nmol):
        n, o = int(b.norb[m]), int(b.mol_mat_off[m])
        Cs = []
        for j in range(3):
            Fm = Fs[j][o:o + n * n].view(n, n)
            Pm = Ps[j][o:o + n * n].view(n, n)
            Cs.append(Fm @ Pm - Pm @ Fm)
        B = torch.tensor([[0.5 * (Cs[i] * Cs[k]).sum().item() for k in range(3)] for i in range(3)],
                         dtype=torch.float64)
        A = torch.zeros(4, 4, dtype=torch.float64)
        A[:3, :3] = B / B[2, 2]
        A[3, :3] = 1.0
        A[:3, 3] = 1.0
        rhs = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=torch.float64)
        cref = torch.linalg.solve(A, rhs)[:3]
        Fref = sum(cref[j].item() * Fs[j][o:o + n * n] for j in range(3))
        got = F[o:o + n * n]
        assert (got - Fref).abs().max().item() < 1e-8 * Fref.abs().max().item(), m


@pytest.mark.parametrize("k", [3, 5, 6])
def test_adaptive_mix_kernel_matches_oracle(k):
    """pyseqm_adaptive_mix (scf_converger [1], scf_loop.py:350-420) against the oracle's
    adaptive_mix on a ragged batch whose molecules need different numbers of renormalisation
    passes (the batch-wide pass count applies to all), at the extrapolating third iterations
    (k = 3, 6; damped at 6) and a damped plain one (k = 5)."""
    from pyseqm_amd import _native as nat
    from pyseqm_amd.engine import SCFEngine
    sp, c = W.alkanes(7, seed=11, nmin=2, nmax=12)
    ob = R.parse(sp, c, "AM1")
    H, T, _, _, _ = R.hcore(ob)
    P0 = R.initial_density(ob)
    _, P1, _ = R.density_eigh(ob, R.fock(ob, P0, H, T))
    _, P2, _ = R.density_eigh(ob, R.fock(ob, 0.5 * (P0 + P1), H, T))
    rng = np.random.default_rng(k)
    n = P0.shape[1]
    old2 = np.einsum("nii->ni", P0) * (1.0 + 0.3 * rng.random((ob.nmol, n)))
    Pprev = 0.5 * (P0 + P1)
    Pref, dprev = R._adaptive_mix(k, Pprev, P2, old2)
    eng = SCFEngine(sp, c, "AM1", DEV)
    b = eng.b
    P = b.from_padded(torch.as_tensor(Pprev, device=DEV))
    Pn = b.from_padded(torch.as_tensor(P2, device=DEV))
    Pold = torch.zeros_like(P)
    # old2 packed per orbital (the padded diagonal's real orbitals, molecule by molecule)
    pm, pr = b.padded_orbital_index()
    o2 = torch.as_tensor(old2[pm, pr], device=DEV)
    orb_off = torch.as_tensor(np.concatenate([[0], np.cumsum(b.norb)[:-1]]).astype(np.int64), device=DEV)
    idx = torch.arange(b.nmol, dtype=torch.int32, device=DEV)
    fac = torch.zeros(b.nmol, dtype=torch.float64, device=DEV)
    mask = torch.zeros(1, dtype=torch.int32, device=DEV)
    Pin = P.clone()
    nat.call("pyseqm_adaptive_mix", b.L, nat.ptr(idx), b.nmol, b.nmax, k, nat.ptr(P), nat.ptr(Pn), nat.ptr(Pold),
             nat.ptr(o2), nat.ptr(orb_off), nat.ptr(fac), nat.ptr(mask), eng.stream)
    torch.cuda.synchronize()
    np.testing.assert_allclose(b.to_padded(P).cpu().numpy(), Pref, atol=1e-12, rtol=0)
    assert torch.equal(Pold, Pin)
    np.testing.assert_allclose(o2.cpu().numpy(), np.einsum("nii->ni", Pprev)[pm, pr], atol=0, rtol=0)
