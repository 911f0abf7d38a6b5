"""The C-ABI library loads and exports every entry point include/pyseqm_hip.h declares (no GPU
compute calls)."""
import ctypes
import os
import re

import pytest

from pyseqm_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pyseqm_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|long long) (pyseqm_\w+)\(", src, re.M)))


def test_header_declares_the_bound_entry_points():
    assert _declared() == sorted(_native.exported_symbols())


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libpyseqm_hip.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    _native.lib()  # argtypes bind without error


def test_custom_ops_registered():
    import torch

    import pyseqm_amd.ops as ops
    for name in ops.OPS:
        assert hasattr(torch.ops.pyseqm, name), name


def test_fock_noalloc_rejects_batches_beyond_its_scratch_before_touching_the_gpu():
    """The size check runs before any HIP call, so it is exercised here without a GPU."""
    lib = _native.lib()
    L = _native.Layout()
    L.nmol, L.natoms = 1, 1000
    L.ntile[0], L.ntile[1], L.ntile[2] = 40000, 0, 0
    need = 160 * 40000 + 40000 + 1000
    assert lib.pyseqm_fock_scratch_doubles(ctypes.byref(L)) == need
    assert need > _native.FOCK_SCRATCH_DOUBLES
    rc = lib.pyseqm_fock_noalloc(ctypes.byref(L), None, None, None, None, None, None, None, None, None, None)
    assert rc == -4
    with pytest.raises(_native.NativeError, match="scratch"):
        _native.call("pyseqm_fock_noalloc", L, None, None, None, None, None, None, None, None, None, None)


def test_fock_scratch_size_matches_header():
    src = open(os.path.join(ROOT, "include", "pyseqm_hip.h")).read()
    m = re.search(r"#define PYSEQM_FOCK_SCRATCH_DOUBLES \((\d+)LL \* 1024 \* 1024\)", src)
    assert m and int(m.group(1)) * 1024 * 1024 == _native.FOCK_SCRATCH_DOUBLES


def test_workgroup_eigensolver_size_rule_matches_the_library():
    from pyseqm_amd.engine import eigh_wg_ok
    lib = _native.lib()
    for n in list(range(1, 40)) + list(range(150, 200)) + [256, 400]:
        for nocc in sorted({0, n // 3, n // 2, (2 * n) // 3, n}):
            assert bool(eigh_wg_ok(n, nocc)) == bool(lib.pyseqm_eigh_density_lds_ok(n, nocc)), (n, nocc)
    assert lib.pyseqm_eigh_density_workspace(7, 100) == 4 * 7 * 100 * 100
