"""Compact binary request format for single-point / XL-BOMD step requests (the inputs of
read_xyz plus method and charge, seqm/seqm_functions/read_xyz.py:16-52): one record per molecule,
records concatenated in a stream.

    offset  size        field
    0       4           magic b"PSQM"
    4       2           format version (1), little-endian
    6       1           method: 0 MNDO, 1 AM1, 2 PM3
    7       1           net charge (signed)
    8       4           natoms (little-endian uint32)
    12      natoms      atomic numbers (uint8), any order
    ...     0-7         zero padding to a multiple of 8 bytes
    ...     24 natoms   coordinates in Angstrom (float64, little-endian, x y z per atom)

Decoding sorts each molecule's atoms by decreasing atomic number (stable), the order the batch
layout requires, exactly as read_xyz does.
"""
import struct

import numpy as np

MAGIC = b"PSQM"
VERSION = 1
METHODS = ("MNDO", "AM1", "PM3")
_HEAD = struct.Struct("<4sHBbI")


def encode(species, coords, method, charge=0):
    """One record: species (natoms,) ints, coords (natoms, 3) Angstrom."""
    Z = np.asarray(species, dtype=np.int64).reshape(-1)
    Z = Z[Z > 0]
    X = np.asarray(coords, dtype=np.float64).reshape(-1, 3)[:len(Z)]
    if Z.size and (Z.max() > 255):
        raise ValueError("atomic numbers above 255 do not fit the record")
    head = _HEAD.pack(MAGIC, VERSION, METHODS.index(method), int(charge), len(Z))
    pad = (-(len(head) + len(Z))) % 8
    return head + Z.astype(np.uint8).tobytes() + b"\0" * pad + X.astype("<f8").tobytes()


def decode(buf):
    """All records of `buf` -> list of (species (natoms,), coords (natoms, 3), method, charge)."""
    out, off, mv = [], 0, memoryview(buf)
    while off < len(buf):
        if len(buf) - off < _HEAD.size:
            raise ValueError(f"truncated record header at byte {off}")
        magic, ver, meth, charge, nat = _HEAD.unpack_from(mv, off)
        if magic != MAGIC or ver != VERSION:
            raise ValueError(f"not a version-{VERSION} request record at byte {off}")
        if meth >= len(METHODS):
            raise ValueError(f"unknown method code {meth} at byte {off}")
        zoff = off + _HEAD.size
        xoff = zoff + nat + (-(_HEAD.size + nat)) % 8
        end = xoff + 24 * nat
        if end > len(buf):
            raise ValueError(f"truncated record at byte {off}")
        Z = np.frombuffer(mv[zoff:zoff + nat], dtype=np.uint8).astype(np.int64)
        X = np.frombuffer(mv[xoff:end], dtype="<f8").reshape(nat, 3).astype(np.float64)
        o = np.argsort(-Z, kind="stable")
        out.append((Z[o], X[o], METHODS[meth], int(charge)))
        off = end
    return out


def batch(records):
    """Padded (species, coords, charges) of decoded records (one method per batch)."""
    K = max((len(r[0]) for r in records), default=0)
    sp = np.zeros((len(records), K), dtype=np.int64)
    xyz = np.zeros((len(records), K, 3))
    for i, (Z, X, _, _) in enumerate(records):
        sp[i, :len(Z)] = Z
        xyz[i, :len(Z)] = X
    return sp, xyz, np.array([r[3] for r in records], dtype=np.int64)
