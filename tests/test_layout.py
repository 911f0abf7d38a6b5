"""Host-side checks of the packed layout (no GPU): pair tiles cover every reference pair exactly
once, and each atom's partial slots cover exactly the pairs it belongs to."""
import numpy as np
import pytest

from oracle import seqm_ref as R
from pyseqm_amd.layout import PackedBatch
from tests.conftest import read_xyz


def _batch(files, method="AM1"):
    sp, c = read_xyz(files)
    return sp, c, PackedBatch(sp, method, "cpu")


def _random_species(rng, nmol, maxat):
    sp = np.zeros((nmol, maxat), dtype=np.int64)
    for m in range(nmol):
        n = rng.integers(2, maxat + 1)
        nh = rng.integers(0, n + 1)
        heavy = np.sort(rng.choice([6, 7, 8], size=nh))[::-1]
        sp[m, :nh] = heavy
        sp[m, nh:n] = 1
    return sp


@pytest.mark.parametrize("seed", [0, 1])
def test_tiles_cover_reference_pairs(seed):
    rng = np.random.default_rng(seed)
    sp = _random_species(rng, 7, 23)
    # electron parity does not matter for the pair layout; make every molecule closed shell
    for m in range(len(sp)):
        if R.TORE[sp[m]].sum() % 2:
            if sp[m, 0] > 1:
                sp[m, 0] = {6: 7, 7: 6, 8: 7}[int(sp[m, 0])]
            else:
                sp[m, np.nonzero(sp[m])[0][-1]] = 0 if (sp[m] > 0).sum() > 2 else 1
                if R.TORE[sp[m]].sum() % 2:
                    sp[m, np.nonzero(sp[m] == 0)[0][0]] = 1
    sp = -np.sort(-sp, axis=1)
    b = PackedBatch(sp, "PM3", "cpu")
    ob = R.parse(sp, np.random.default_rng(1).normal(size=sp.shape + (3,)) * 3, "PM3")
    ref = b.pair_ref_host[b.lane_valid]
    assert len(ref) == b.npairs == len(ob.idxi)
    assert np.array_equal(np.sort(ref), np.arange(b.npairs))
    # lane (tile, r, c) -> atoms; compare with the oracle's reference pair list
    T = b.tiles_host
    lanes = np.nonzero(b.lane_valid)[0]
    t, r, c = lanes // 64, (lanes % 64) // 8, lanes % 8
    i, j = T[t, 1] + r, T[t, 2] + c
    order = np.argsort(b.pair_ref_host[lanes])
    assert np.array_equal(i[order], ob.idxi) and np.array_equal(j[order], ob.idxj)
    # slots of atom a: pairs counted through its row/column slots == pairs containing a
    slot_atom = {}
    for a in range(b.natoms):
        for s in b.csr_idx_host[b.csr_off_host[a]:b.csr_off_host[a + 1]]:
            slot_atom[int(s)] = a
    cnt = np.zeros(b.natoms, dtype=np.int64)
    for ln in lanes:
        tt, rr, cc = ln // 64, (ln % 64) // 8, ln % 8
        assert slot_atom[tt * 16 + rr] == T[tt, 1] + rr
        assert slot_atom[tt * 16 + 8 + cc] == T[tt, 2] + cc
        cnt[T[tt, 1] + rr] += 1
        cnt[T[tt, 2] + cc] += 1
    expect = np.bincount(np.concatenate([ob.idxi, ob.idxj]), minlength=b.natoms)
    assert np.array_equal(cnt, expect)


def test_orbital_layout_matches_reference_pack_order():
    sp, c, b = _batch(["methanal.1.xyz", "benzene.xyz"])
    ob = R.parse(sp, c, "AM1")
    om, orow = b.padded_orbital_index()
    for m in range(b.nmol):
        assert np.array_equal(orow[om == m], R.orbital_index(ob, m))
    assert np.array_equal(b.norb, ob.norb) and np.array_equal(b.nocc, ob.nocc)


def test_parser_errors_match_reference():
    sp, c = read_xyz(["methane.xyz"])
    with pytest.raises(ValueError, match="non-increasing"):
        PackedBatch(sp[:, ::-1].copy(), "AM1", "cpu")
    with pytest.raises(ValueError, match="even number of electrons"):
        PackedBatch(sp, "AM1", "cpu", charges=[1])
    with pytest.raises(ValueError, match="not supported"):
        PackedBatch(np.array([[35, 1]]), "AM1", "cpu")


def test_pack_unpack_match_orbital_selection():
    """pack keeps exactly the real orbitals of each molecule in reference order (heavy blocks,
    then hydrogen s rows); unpack is its inverse on real rows and zero elsewhere."""
    import torch
    from oracle import seqm_ref as R
    from pyseqm_amd.seqm_functions.pack import pack, unpack
    sp, c = read_xyz(["methanal.1.xyz", "benzene.xyz", "methane.xyz"])
    ob = R.parse(sp, c, "AM1")
    rng = np.random.default_rng(0)
    n = 4 * ob.molsize
    X = rng.normal(size=(ob.nmol, n, n))
    Xp = pack(torch.as_tensor(X), torch.as_tensor(ob.nHeavy), torch.as_tensor(ob.nHydro)).numpy()
    N = int(ob.norb.max())
    assert Xp.shape == (ob.nmol, N, N)
    for m in range(ob.nmol):
        idx = R.orbital_index(ob, m)
        k = len(idx)
        np.testing.assert_array_equal(Xp[m, :k, :k], X[m][np.ix_(idx, idx)])
        assert not Xp[m, k:].any() and not Xp[m, :, k:].any()
    Xu = unpack(torch.as_tensor(Xp), torch.as_tensor(ob.nHeavy), torch.as_tensor(ob.nHydro), n).numpy()
    for m in range(ob.nmol):
        idx = R.orbital_index(ob, m)
        np.testing.assert_array_equal(Xu[m][np.ix_(idx, idx)], X[m][np.ix_(idx, idx)])
        mask = np.ones((n, n), bool)
        mask[np.ix_(idx, idx)] = False
        assert not Xu[m][mask].any()


def test_vectorised_tile_regions_match_per_molecule_construction():
    from pyseqm_amd.layout import _region, _regions
    rng = np.random.default_rng(4)
    nmol = 40
    nh = rng.integers(0, 30, nmol)
    nH = rng.integers(0, 40, nmol)
    a0 = np.concatenate([[0], np.cumsum(nh + nH)])[:-1]
    zero = np.zeros(nmol, np.int64)
    for args, tri in (((zero, nh, zero, nh), True), ((zero, nh, nh, nH), False), ((nh, nH, nh, nH), True)):
        t, cnt = _regions(a0, *args, tri)
        ref = [_region(m, a0[m], args[0][m], args[1][m], args[2][m], args[3][m], tri) for m in range(nmol)]
        np.testing.assert_array_equal(t, np.concatenate(ref))
        np.testing.assert_array_equal(cnt, [len(r) for r in ref])
