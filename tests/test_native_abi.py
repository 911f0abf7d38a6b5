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
    return sorted(set(re.findall(r"^int (pyseqm_\w+)\(", src, re.M)))


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
