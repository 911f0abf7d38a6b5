"""GPU (gfx950) kernels vs the FP64 oracle and the reference goldens.

Bars: integrals / Fock / energies of one geometry to 1e-10 relative to the matrix scale
(FP64 summation-order differences only); converged SCF energies and heats of formation within
1e-6 eV of the oracle (tighter, 1e-9 eV, against the reference goldens at the goldens' own
settings); the kernels are deterministic (bitwise-identical reruns).
"""
import numpy as np
import pytest
import torch

from oracle import seqm_ref as R
from pyseqm_amd import workloads as W
from pyseqm_amd.engine import SCFEngine
from tests.conftest import golden, read_xyz

pytestmark = pytest.mark.gpu
DEV = "cuda"

# small third-row molecules (S: jcall 431 / 5 / 6 overlaps), Angstrom
H2S = (np.array([[16, 1, 1]]), np.array([[[0.0, 0.0, 0.1], [0.96, 0.0, 0.93], [-0.97, 0.05, 0.92]]]))
CH3SH = (np.array([[16, 6, 1, 1, 1, 1]]),
         np.array([[[0.0, 0.0, 0.0], [1.82, 0.0, 0.05], [2.19, 1.03, 0.0], [2.18, -0.53, 0.89],
                    [2.2, -0.5, -0.9], [-0.3, 1.3, 0.1]]]))
HSSH = (np.array([[16, 16, 1, 1]]),
        np.array([[[0.0, 0.0, 0.0], [2.06, 0.0, 0.0], [-0.3, 1.32, 0.1], [2.35, 0.1, 1.31]]]))


def _cat(*mols):
    K = max(m[0].shape[1] for m in mols)
    sp = np.zeros((len(mols), K), dtype=np.int64)
    xyz = np.zeros((len(mols), K, 3))
    for i, (s, x) in enumerate(mols):
        sp[i, :s.shape[1]] = s[0]
        xyz[i, :s.shape[1]] = x[0]
    return sp, xyz


def _inputs(name):
    if name == "methanal+benzene":
        return read_xyz(["methanal.1.xyz", "benzene.xyz"])
    if name == "methane+benzene":
        return read_xyz(["methane.xyz", "benzene.xyz"])
    if name == "sulfur":
        return _cat(H2S, CH3SH, HSSH)
    if name == "water8":
        return W.water_clusters(8, 2, seed=3)
    if name == "alkanes":
        return W.alkanes(6, seed=5, nmax=9)
    if name == "benzene4":
        return W.benzene(4, seed=2)
    if name == "tilekinds":
        return _tile_kind_batch()
    raise KeyError(name)


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1.0)
    err = np.abs(a - b).max()
    assert err <= rtol * scale, (err, scale)


@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
@pytest.mark.parametrize("name", ["methanal+benzene", "sulfur", "water8", "alkanes"])
def test_integrals_and_fock_match_oracle(method, name):
    sp, c = _inputs(name)
    eng = SCFEngine(sp, c, method, DEV)
    ob = R.parse(sp, c, method)
    H, T, di, gam, mp = R.hcore(ob)
    _close(eng.w_reference().cpu().numpy(), R.tensor_to_w(T), 1e-12)
    _close(eng.b.to_padded(eng.H).cpu().numpy(), H, 1e-12)
    eo = np.zeros(ob.nmol)
    np.add.at(eo, ob.pair_mol, R.pair_nuclear_energy(ob, gam))
    _close(eng.enuc.cpu().numpy(), eo, 1e-12)
    # Fock of a random symmetric density (exercises every J/K term), and of the initial guess
    rng = np.random.default_rng(0)
    Pr = rng.normal(size=H.shape)
    Pr = Pr + Pr.transpose(0, 2, 1)
    for Po in (R.initial_density(ob), Pr):
        for a in np.nonzero(ob.Z == 1)[0]:  # hydrogens carry only an s orbital
            o = 4 * ob.atom_loc[a]
            Po[ob.atom_mol[a], o + 1:o + 4, :] = 0.0
            Po[ob.atom_mol[a], :, o + 1:o + 4] = 0.0
        P = eng.b.from_padded(torch.as_tensor(Po, device=DEV))
        F = torch.empty_like(P)
        ee = eng.fock(P, F)
        Fo = R.fock(ob, Po, H, T)
        _close(eng.b.to_padded(F).cpu().numpy(), Fo, 1e-12)
        _close(ee.cpu().numpy(), R.elec_energy(Po, Fo, H), 1e-12)


@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
def test_methane_golden_on_gpu(method):
    sp, c = read_xyz(["methane.xyz"])
    eng = SCFEngine(sp, c, method, DEV)
    r = eng.scf(eps=1e-6, converger=(1,))
    en = eng.energies(r.P)
    g = golden(f"smoke_single_point_{method}")
    for k in ("Etot", "Eelec", "Enuc"):
        assert abs(en[k].item() - g[k]) < 1e-9, (k, en[k].item(), g[k])


def test_batch_golden_on_gpu():
    sp, c = read_xyz(["methane.xyz", "benzene.xyz"])
    eng = SCFEngine(sp, c, "AM1", DEV)
    r = eng.scf(eps=1e-6, converger=(1,))
    en = eng.energies(r.P)
    np.testing.assert_allclose(en["Etot"].cpu().numpy(), golden("batch_single_point_am1")["Etot"], atol=1e-9, rtol=0)


@pytest.mark.parametrize("name,method", [("sulfur", "PM3"), ("water8", "PM3"), ("alkanes", "AM1"),
                                         ("methanal+benzene", "MNDO")])
@pytest.mark.parametrize("conv", [(2,), (0, 0.5)])
def test_scf_matches_oracle(name, method, conv):
    sp, c = _inputs(name)
    eng = SCFEngine(sp, c, method, DEV)
    r = eng.scf(eps=1e-8, converger=conv)
    en = eng.energies(r.P)
    ro = R.single_point(sp, c, method, eps=1e-8, converger=conv)
    assert not r.notconverged.any() and not ro["notconverged"].any()
    for k in ("Hf", "Etot", "Eelec", "Enuc"):
        np.testing.assert_allclose(en[k].cpu().numpy(), ro[k], atol=1e-6, rtol=0)
    if conv[0] == 2:
        assert abs(r.iterations - ro["iterations"]) <= 1


def test_c1_benzene_batch_pm3_pulay():
    sp, c = W.benzene(16, seed=0)
    eng = SCFEngine(sp, c, "PM3", DEV)
    r = eng.scf(eps=1e-8, converger=(2,))
    en = eng.energies(r.P)
    ro = R.single_point(sp, c, "PM3", eps=1e-8, converger=(2,))
    np.testing.assert_allclose(en["Hf"].cpu().numpy(), ro["Hf"], atol=1e-6, rtol=0)
    e, _ = eng.eig(en["F"])
    np.testing.assert_allclose(e.cpu().numpy(), ro["e"], atol=1e-6, rtol=0)


@pytest.mark.parametrize("name", ["methanal+benzene", "water8", "benzene4"])
def test_fock_per_molecule_kernel_matches_tile_kernels(name):
    """Batches whose molecules have at most PYSEQM_FOCK_MOL_ATOMS atoms take the per-molecule
    Fock kernel; the same build through the tile-parallel kernels (taken when the layout gives no
    molecule size) agrees to 1e-12, and the per-molecule build is bitwise repeatable."""
    sp, c = _inputs(name)
    eng = SCFEngine(sp, c, "PM3", DEV)
    L = eng.b.L
    assert 0 < L.max_mol_atoms <= 24
    g = torch.Generator(device=DEV).manual_seed(4)
    P = eng.initial_density()
    P = P + 0.05 * torch.randn(P.shape, generator=g, device=DEV, dtype=torch.float64)
    F1, F2, F3 = (torch.full_like(P, float("nan")) for _ in range(3))
    e1 = eng.fock(P, F1).clone()
    e3 = eng.fock(P, F3).clone()
    saved = L.max_mol_atoms
    L.max_mol_atoms = 0
    try:
        e2 = eng.fock(P, F2).clone()
    finally:
        L.max_mol_atoms = saved
    assert torch.equal(F1, F3) and torch.equal(e1, e3)
    assert not torch.isnan(F1).any()
    scale = float(F2.abs().max())
    assert float((F1 - F2).abs().max()) <= 1e-12 * scale
    assert float((e1 - e2).abs().max()) <= 1e-12 * float(e2.abs().max())


def test_fock_per_molecule_kernel_refuses_understated_layouts():
    """A layout whose max_mol_atoms understates its molecules sends them to the per-molecule
    kernel, which checks each molecule against its LDS arrays: larger ones are skipped with a NaN
    Eelec (nothing written out of bounds), molecules that fit are built as usual."""
    sp, c = W.alkanes(24, seed=7, nmin=2, nmax=12)
    eng = SCFEngine(sp, c, "AM1", DEV)
    L = eng.b.L
    P = eng.initial_density()
    F0 = torch.empty_like(P)
    e0 = eng.fock(P, F0).clone()
    saved = L.max_mol_atoms
    L.max_mol_atoms = 20
    try:
        e1 = eng.fock(P, torch.empty_like(P)).clone()
    finally:
        L.max_mol_atoms = saved
    big = torch.as_tensor(eng.b.nat > 24, device=DEV)
    assert big.any() and (~big).any()
    assert torch.isnan(e1[big]).all()
    torch.testing.assert_close(e1[~big], e0[~big], rtol=1e-12, atol=1e-9)


def test_fock_is_deterministic():
    sp, c = W.alkanes(32, seed=2)
    eng = SCFEngine(sp, c, "AM1", DEV)
    P = eng.initial_density()
    F1, F2 = torch.empty_like(P), torch.empty_like(P)
    e1 = eng.fock(P, F1).clone()
    e2 = eng.fock(P, F2).clone()
    assert torch.equal(F1, F2) and torch.equal(e1, e2)
    H1 = eng.H.clone()
    eng.set_coordinates(torch.as_tensor(c))
    assert torch.equal(H1, eng.H)


def test_molecule_alone_equals_in_batch():
    sp, c = W.alkanes(12, seed=7)
    eng = SCFEngine(sp, c, "AM1", DEV)
    hf = eng.energies(eng.scf(eps=1e-8).P)["Hf"].cpu().numpy()
    for m in (0, 5, 11):
        e1 = SCFEngine(sp[m:m + 1], c[m:m + 1], "AM1", DEV)
        h1 = e1.energies(e1.scf(eps=1e-8).P)["Hf"].cpu().numpy()
        assert abs(h1[0] - hf[m]) < 1e-7


@pytest.mark.parametrize("files,method,eps,gname", [
    (["methane.xyz"], "MNDO", 1e-6, "smoke_single_point_MNDO"),
    (["methane.xyz"], "AM1", 1e-6, "smoke_single_point_AM1"),
    (["methane.xyz"], "PM3", 1e-6, "smoke_single_point_PM3"),
    (["methanal.1.xyz", "methanal.2.xyz", "methanal.3.xyz"], "AM1", 1e-7, "ground_force_batch_methanal"),
    (["methane.xyz", "benzene.xyz"], "AM1", 1e-7, "ground_force_batch_mixed"),
])
def test_forces_match_goldens(files, method, eps, gname):
    sp, c = read_xyz(files)
    eng = SCFEngine(sp, c, method, DEV)
    r = eng.scf(eps=eps, converger=(1,))
    f = eng.b.atom_to_padded(eng.forces(r.P)).cpu().numpy()
    g = np.array(golden(gname)["force"]).reshape(f.shape)
    np.testing.assert_allclose(f, g, atol=1e-6, rtol=0)


@pytest.mark.parametrize("name,method", [("sulfur", "PM3"), ("water8", "MNDO"), ("tilekinds", "AM1")])
def test_forces_match_oracle(name, method):
    sp, c = _inputs(name)
    eng = SCFEngine(sp, c, method, DEV)
    r = eng.scf(eps=1e-8, converger=(2,))
    f = eng.b.atom_to_padded(eng.forces(r.P)).cpu().numpy()
    Po = eng.b.to_padded(r.P).cpu().numpy()
    fo = R.forces_fd(sp, c, method, Po)
    np.testing.assert_allclose(f, fo, atol=1e-6, rtol=0)


def test_diis_solve_matches_eigh_pseudo_inverse():
    """The Jacobi DIIS solve against the reference's eigh-based pseudo-inverse (scf_loop.py:1015-1031)
    on bordered B matrices of nearly dependent residuals: coefficients to 1e-8 relative where the
    system is kept (cond <= 1e7), identical reset decision."""
    from pyseqm_amd import _native as nat
    sp, c = _inputs("alkanes")
    eng = SCFEngine(sp, c, "AM1", DEV)
    rng = np.random.default_rng(9)
    nf = 10
    nmol = eng.b.nmol
    for cF in (2, 5, 10):
        E = np.zeros((nmol, nf + 1, nf + 1))
        for m in range(nmol):
            X = rng.normal(size=(cF, 40))
            X[1:] = X[0] + 10.0 ** rng.uniform(-4, 0) * X[1:]
            B = X @ X.T
            Em = -np.ones((cF + 1, cF + 1))
            Em[:cF, :cF] = B
            Em[cF, cF] = 0.0
            E[m, :cF + 1, :cF + 1] = np.tril(Em)
            E[m] += np.triu(rng.normal(size=(nf + 1, nf + 1)), 1)  # stale upper entries are ignored
        slot = cF - 1
        Et = torch.as_tensor(E, device=DEV)
        coef = torch.zeros(nmol, nf, dtype=torch.float64, device=DEV)
        reset = torch.zeros(nmol, dtype=torch.uint8, device=DEV)
        idx = torch.arange(nmol, dtype=torch.int32, device=DEV)
        nat.call("pyseqm_diis_solve", eng.b.L, nat.ptr(idx), nmol, nf, cF, slot, nat.ptr(Et), nat.ptr(coef),
                 nat.ptr(reset), eng.stream)
        for m in range(nmol):
            EV = np.tril(E[m, :cF + 1, :cF + 1])
            EV = EV + np.tril(EV, -1).T
            EV[:cF, :cF] /= max(EV[slot, slot], 1e-15)
            w, Q = np.linalg.eigh(EV)
            inv = np.where(np.abs(w) > 1e-13, 1.0 / np.where(np.abs(w) > 1e-13, w, 1.0), 0.0)
            cref = -(Q[:cF] * inv[None] * Q[cF][None]).sum(1)
            cond = np.abs(w).max() / np.abs(w).min()
            assert bool(reset[m].item()) == (cond > 1e7), (m, cond)
            if cond <= 1e7:
                np.testing.assert_allclose(coef[m, :cF].cpu().numpy(), cref, rtol=0,
                                           atol=1e-8 * max(1.0, np.abs(cref).max()))


def _tile_kind_batch():
    """H2 (H-H pairs only), HF (heavy-H only), N2 (heavy-heavy only) and H2O, padded to 3 atoms."""
    sp = np.array([[1, 1, 0], [9, 1, 0], [7, 7, 0], [8, 1, 1]])
    xyz = np.zeros((4, 3, 3))
    xyz[0, :2] = [[0.0, 0.0, 0.0], [0.74, 0.02, 0.0]]
    xyz[1, :2] = [[0.0, 0.0, 0.0], [0.0, 0.03, 0.92]]
    xyz[2, :2] = [[0.0, 0.0, 0.0], [1.10, 0.0, 0.01]]
    xyz[3] = [[0.0, 0.0, 0.0], [0.96, 0.0, 0.0], [-0.24, 0.93, 0.0]]
    return sp, xyz


@pytest.mark.parametrize("method", ["MNDO", "AM1", "PM3"])
def test_single_tile_kind_molecules_match_oracle(method):
    """Molecules whose pairs are all of one tile kind, in one ragged batch and alone (H2 alone and N2
    alone are batches with no tiles at all of two of the three kinds): SCF energies within 1e-6 eV of
    the oracle, each molecule alone equal to its batch result, and the mixed-precision density step
    within 1e-4 eV/atom."""
    sp, c = _tile_kind_batch()
    eng = SCFEngine(sp, c, method, DEV)
    r = eng.scf(eps=1e-8, converger=(2,))
    en = eng.energies(r.P)
    ro = R.single_point(sp, c, method, eps=1e-8, converger=(2,))
    assert not r.notconverged.any() and not ro["notconverged"].any()
    for k in ("Hf", "Etot", "Eelec", "Enuc"):
        np.testing.assert_allclose(en[k].cpu().numpy(), ro[k], atol=1e-6, rtol=0)
    hf = en["Hf"].cpu().numpy()
    for m in (0, 2):
        e1 = SCFEngine(sp[m:m + 1, :2], c[m:m + 1, :2], method, DEV)
        h1 = e1.energies(e1.scf(eps=1e-8).P)["Hf"].cpu().numpy()
        assert abs(h1[0] - hf[m]) < 1e-7, (m, h1[0], hf[m])
    e2 = SCFEngine(sp, c, method, DEV)
    r2 = e2.scf(eps=1e-5, converger=(2,), density="sp2")
    assert not r2.notconverged.any()
    h2 = e2.energies(r2.P)["Hf"].cpu().numpy()
    nat = (sp > 0).sum(1)
    assert np.all(np.abs(h2 - ro["Hf"]) / nat < 1e-4), (h2 - ro["Hf"]) / nat
