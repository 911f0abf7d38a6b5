"""The device tile builder (pyseqm_build_tiles) against the host enumeration: the same tile records,
integral offsets, buffer sizes and per-atom slot lists, entry for entry, on ragged batches that
include molecules of one tile kind and triangles whose last diagonal tile holds a single atom."""
import numpy as np
import pytest

from pyseqm_amd import workloads as W
from pyseqm_amd.layout import PackedBatch

pytestmark = pytest.mark.gpu


def _cat(*batches):
    K = max(b[0].shape[1] for b in batches)
    rows = [np.pad(s, ((0, 0), (0, K - s.shape[1]))) for s, _ in batches]
    return np.concatenate(rows)


def _odd():
    # 9 heavy atoms and 17 hydrogens (last diagonal tiles with one atom), 1 heavy + 1 H, H2, N2, 8 + 8
    rows = [[6] * 9 + [1] * 17, [9, 1], [1, 1], [7, 7], [6] * 8 + [1] * 8, [8] + [1] * 2, [6] * 17 + [1] * 9]
    K = max(map(len, rows))
    return np.array([r + [0] * (K - len(r)) for r in rows])


@pytest.mark.parametrize("name", ["alkanes", "waters", "odd", "mixed"])
def test_device_tiles_equal_host_tiles(name):
    if name == "alkanes":
        sp, method = W.alkanes(1024, seed=0)[0], "AM1"
    elif name == "waters":
        sp, method = W.water_clusters(256, 2, seed=0)[0], "PM3"
    elif name == "odd":
        sp, method = _odd(), "AM1"
    else:
        sp, method = _cat((_odd(), None), (W.alkanes(40, seed=3)[0], None), (W.water_clusters(20, 3, seed=1)[0], None)), "AM1"
    from pyseqm_amd.seqm_functions.constants import TORE
    charges = np.asarray(TORE)[sp].sum(1).astype(np.int64) % 2  # closed shells (odd counts made cations)
    dev = PackedBatch(sp, method, "cuda", charges=charges)
    host = PackedBatch(sp, method, "cpu", charges=charges)
    assert dev.ntile == host.ntile and dev.w_size == host.w_size
    np.testing.assert_array_equal(dev.mol_tile_off, host.mol_tile_off)
    np.testing.assert_array_equal(dev.tiles_host, host.tiles_host)
    np.testing.assert_array_equal(dev.tile_woff_host, host.tile_woff_host)
    np.testing.assert_array_equal(dev.tile_wstride, host.tile_wstride)
    np.testing.assert_array_equal(dev.csr_off_host, host.csr_off_host)
    np.testing.assert_array_equal(dev.csr_idx_host, host.csr_idx_host)
    assert dev.nslots == len(host.csr_idx_host)
