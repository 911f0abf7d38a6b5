"""Packed, padding-free batch layout on the device.

Builds, from the reference's padded ``species`` (nmol, molsize) input, everything the gfx950
kernels index with: per-molecule atom / orbital / matrix offsets, the 8x8 pair tiles of the three
pair kinds (heavy-heavy, heavy-H, H-H) and the per-atom lists of tile partial slots.  Parsing
follows the reference Parser (seqm/basics.py:219-403): atoms are the non-zero entries of each
row in order, pairs are i < j within a molecule in row-major order (the reference pair index
is kept in ``pair_ref`` for outputs in reference order), nocc = sum(tore)/2 - charge/2.
"""
import ctypes

import numpy as np
import torch

from . import _native
from .seqm_functions.constants import QN, TORE, MAX_Z
from .seqm_functions.parameters import supported_elements

TILE = 8
W_PER_LANE = (100, 10, 1)  # integrals stored per atom pair: heavy-heavy, heavy-H, H-H


def check_species(species):
    """Rows must be non-increasing (seqm/Molecule.py:188-206); same message as the reference."""
    ok = np.all(species[:, :-1] >= species[:, 1:], axis=1)
    if not ok.all():
        bad = np.nonzero(~ok)[0].tolist()
        rows = ", ".join(map(str, bad))
        word, verb = ("row", "is") if len(bad) == 1 else ("rows", "are")
        raise ValueError(f"species must be non-increasing along each row, but {word} {rows} {verb} not sorted.")


def _region(mol, a0, r_start, nr, c_start, nc, tri):
    """Tiles covering rows [r_start, r_start+nr) x cols [c_start, c_start+nc) of molecule `mol`
    (local atom indices); tri: upper triangle i < j of a square region."""
    nbr, nbc = -(-nr // TILE), -(-nc // TILE)
    if nbr == 0 or nbc == 0:
        return np.zeros((0, 6), dtype=np.int64)
    bi, bj = np.meshgrid(np.arange(nbr), np.arange(nbc), indexing="ij")
    bi, bj = bi.ravel(), bj.ravel()
    if tri:
        keep = bj >= bi
        bi, bj = bi[keep], bj[keep]
    nrow = np.minimum(TILE, nr - TILE * bi)
    ncol = np.minimum(TILE, nc - TILE * bj)
    diag = (bi == bj) & tri
    keep = ~(diag & (nrow < 2))  # a diagonal tile with one atom has no pair
    t = np.stack([np.full(len(bi), mol), a0 + r_start + TILE * bi, a0 + c_start + TILE * bj, nrow, ncol,
                  diag.astype(np.int64)], 1)
    return t[keep]


def _regions(a0, r_start, nr, c_start, nc, tri):
    """_region for every molecule at once (same tiles, same molecule-major / row-major order).
    Arrays are per molecule; returns (tiles (N, 6), tiles per molecule)."""
    nmol = len(nr)
    nbr, nbc = -(-nr // TILE), -(-nc // TILE)
    cnt = nbr * nbc
    mol = np.repeat(np.arange(nmol), cnt)
    k = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    bi = k // np.maximum(nbc[mol], 1)
    bj = k % np.maximum(nbc[mol], 1)
    keep = (bj >= bi) if tri else np.ones(len(k), bool)
    mol, bi, bj = mol[keep], bi[keep], bj[keep]
    nrow = np.minimum(TILE, nr[mol] - TILE * bi)
    ncol = np.minimum(TILE, nc[mol] - TILE * bj)
    diag = (bi == bj) & tri
    keep = ~(diag & (nrow < 2))
    t = np.stack([mol, a0[mol] + r_start[mol] + TILE * bi, a0[mol] + c_start[mol] + TILE * bj, nrow, ncol,
                  diag.astype(np.int64)], 1)[keep]
    return t.astype(np.int64), np.bincount(mol[keep], minlength=nmol)


class MatrixLayout:
    """Layout of nmol dense n x n blocks stored back to back (a (nmol, n, n) tensor), for the
    matrix-only kernels (density steps, energies) on inputs given in the reference's dense
    layouts.  norb[m] = n for every block; occupations per block."""

    def __init__(self, nmol, n, nocc, device):
        self.nmol, self.n, self.device = int(nmol), int(n), torch.device(device)
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.int32, device=self.device)
        nocc = np.broadcast_to(np.asarray(torch.as_tensor(nocc).cpu(), dtype=np.int64).reshape(-1), (self.nmol,))
        self.norb = np.full(self.nmol, self.n, dtype=np.int64)
        self.nocc = nocc.copy()
        self.d = {"mol_norb": i32(self.norb), "mol_nocc": i32(self.nocc),
                  "mol_mat_off": torch.arange(self.nmol, dtype=torch.int64, device=self.device) * (self.n * self.n),
                  "mol_atom_off": torch.zeros(self.nmol + 1, dtype=torch.int32, device=self.device)}
        L = _native.Layout()
        L.nmol, L.natoms, L.npairs = self.nmol, 0, 0
        for k in range(3):
            L.ntile[k] = 0
        for name in ("mol_norb", "mol_nocc", "mol_mat_off", "mol_atom_off"):
            setattr(L, name, self.d[name].data_ptr())
        self.L = L
        self.total = self.nmol * self.n * self.n
        self.nmax = self.n


class PackedBatch:
    """Host description + device tensors of one packed batch."""

    def __init__(self, species, method, device, charges=None):
        sp = species.detach().cpu().numpy() if torch.is_tensor(species) else np.asarray(species)
        sp = sp.astype(np.int64)
        if sp.ndim == 1:
            sp = sp[None]
        check_species(sp)
        self.method = method
        self.device = torch.device(device)
        self.species = sp
        nmol, molsize = sp.shape
        self.nmol, self.molsize = nmol, molsize
        real = sp > 0
        flat = np.nonzero(real.reshape(-1))[0]
        self.real_flat = flat                      # padded flat index of each atom
        self.Z = sp.reshape(-1)[flat]
        self.natoms = len(flat)
        if self.natoms and (self.Z.max() > MAX_Z or np.any(~np.isin(self.Z, supported_elements(method)))):
            raise ValueError("\nError from diat.py, overlap matrix\nSome elements are not supported yet")
        self.atom_mol = flat // molsize
        self.atom_loc = flat % molsize
        self.nat = real.sum(1)
        self.nheavy = (sp > 1).sum(1)
        self.nhydro = (sp == 1).sum(1)
        tore = np.asarray(TORE, dtype=np.float64)
        nel = tore[sp].sum(1).astype(np.int64)
        if charges is not None:
            ch = charges.detach().cpu().numpy() if torch.is_tensor(charges) else np.asarray(charges)
            nel = nel - np.broadcast_to(ch, (nmol,)).astype(np.int64)
        if np.any(nel % 2 == 1):
            raise ValueError("RHF setting requires closed shell systems (even number of electrons)")
        self.nocc = nel // 2
        self.norb = 4 * self.nheavy + self.nhydro
        self.mol_atom_off = np.concatenate([[0], np.cumsum(self.nat)]).astype(np.int64)
        n2 = self.norb.astype(np.int64) ** 2
        self.mol_mat_off = np.concatenate([[0], np.cumsum(n2)])[:-1].astype(np.int64)
        self.total = int(n2.sum())
        self.nmax = int(self.norb.max()) if nmol else 0
        loc = self.atom_loc
        nh_of = self.nheavy[self.atom_mol]
        self.atom_ao = np.where(self.Z > 1, 4 * loc, 4 * nh_of + (loc - nh_of)).astype(np.int64)
        self.npairs_mol = self.nat * (self.nat - 1) // 2
        self.pair_base = np.concatenate([[0], np.cumsum(self.npairs_mol)]).astype(np.int64)
        self.npairs = int(self.pair_base[-1])
        if self.device.type == "cuda":
            self._tile_counts()
            self._to_device(device_tiles=True)
        else:
            self._build_tiles()
            self._to_device()

    # ------------------------------------------------------------------ tiles
    def _tile_counts(self):
        """Per-molecule tile and slot counts in closed form (O(nmol)); the tiles themselves are built
        on the device (pyseqm_build_tiles).  A triangle region of n atoms has nb (nb + 1) / 2 tiles of
        nb = ceil(n / 8) per side, less a last diagonal tile holding one atom, and (nb + 1) n partial
        slots (less that tile's 2); a rectangle nr x nc has nbr nbc tiles and nbc nr + nbr nc slots."""
        nh, nH = self.nheavy.astype(np.int64), self.nhydro.astype(np.int64)
        nbx, nbh = -(-nh // TILE), -(-nH // TILE)
        skip_x, skip_h = (nh % TILE == 1).astype(np.int64), (nH % TILE == 1).astype(np.int64)
        counts = [nbx * (nbx + 1) // 2 - skip_x, nbx * nbh, nbh * (nbh + 1) // 2 - skip_h]
        self.ntile = [int(c.sum()) for c in counts]
        base = np.cumsum([0] + self.ntile)
        self.tile_base = base[:3]
        self.mol_tile_off = np.concatenate([np.concatenate([[0], np.cumsum(c)]) + b
                                            for c, b in zip(counts, base[:3])]).astype(np.int64)
        self.tile_kind = np.concatenate([np.full(n, k) for k, n in enumerate(self.ntile)]).astype(np.int64)
        slots = ((nbx + 1) * nh - 2 * skip_x) + (nbh * nh + nbx * nH) + ((nbh + 1) * nH - 2 * skip_h)
        self.nslots = int(slots.sum())
        self._host_tiles = None
        self._lanes = None

    def _fetch_tiles(self):
        """Host copies of the device-built tile records, integral offsets and slot lists."""
        if self._host_tiles is None:
            nt = sum(self.ntile)
            rec = self.d["tiles"][:4 * nt].cpu().numpy().astype(np.int64).reshape(-1, 4)
            T = np.zeros((nt, 6), dtype=np.int64)
            T[:, :3] = rec[:, :3]
            T[:, 3] = rec[:, 3] & 0xff
            T[:, 4] = (rec[:, 3] >> 8) & 0xff
            T[:, 5] = (rec[:, 3] >> 16) & 1
            nrow, ncol, diag = T[:, 3], T[:, 4], T[:, 5] != 0
            nvalid = np.where(diag, nrow * (nrow - 1) // 2, nrow * ncol).astype(np.int64)
            self._host_tiles = (T, self.d["tile_woff"][:nt].cpu().numpy().astype(np.int64), (nvalid + 15) // 16 * 16,
                                self.d["csr_off"].cpu().numpy().astype(np.int64),
                                self.d["csr_idx"][:self.nslots].cpu().numpy().astype(np.int64))
        return self._host_tiles

    @property
    def tiles_host(self):
        return self._tiles_host if hasattr(self, "_tiles_host") else self._fetch_tiles()[0]

    @tiles_host.setter
    def tiles_host(self, v):
        self._tiles_host = v

    @property
    def tile_woff_host(self):
        return self._tile_woff_host if hasattr(self, "_tile_woff_host") else self._fetch_tiles()[1]

    @tile_woff_host.setter
    def tile_woff_host(self, v):
        self._tile_woff_host = v

    @property
    def tile_wstride(self):
        return self._tile_wstride if hasattr(self, "_tile_wstride") else self._fetch_tiles()[2]

    @tile_wstride.setter
    def tile_wstride(self, v):
        self._tile_wstride = v

    @property
    def csr_off_host(self):
        return self._csr_off_host if hasattr(self, "_csr_off_host") else self._fetch_tiles()[3]

    @csr_off_host.setter
    def csr_off_host(self, v):
        self._csr_off_host = v

    @property
    def csr_idx_host(self):
        return self._csr_idx_host if hasattr(self, "_csr_idx_host") else self._fetch_tiles()[4]

    @csr_idx_host.setter
    def csr_idx_host(self, v):
        self._csr_idx_host = v

    def _build_tiles(self):
        kinds = []
        a0 = self.mol_atom_off[:-1].astype(np.int64)
        nh, nH = self.nheavy.astype(np.int64), self.nhydro.astype(np.int64)
        zero = np.zeros_like(nh)
        for kind in range(3):
            if kind == 0:
                tiles, counts = _regions(a0, zero, nh, zero, nh, True)
            elif kind == 1:
                tiles, counts = _regions(a0, zero, nh, nh, nH, False)
            else:
                tiles, counts = _regions(a0, nh, nH, nh, nH, True)
            kinds.append((tiles, np.concatenate([[0], np.cumsum(counts)])))
        self.ntile = [len(k[0]) for k in kinds]
        base = np.cumsum([0] + self.ntile)
        self.tile_base = base[:3]
        T = np.concatenate([k[0] for k in kinds]) if sum(self.ntile) else np.zeros((0, 6), np.int64)
        self.tiles_host = T
        self.mol_tile_off = np.concatenate([k[1] + b for k, b in zip(kinds, base[:3])]).astype(np.int64)
        nt = len(T)
        self.tile_kind = np.concatenate([np.full(n, k) for k, n in enumerate(self.ntile)]).astype(np.int64)
        # valid lanes per tile in closed form: nrow x ncol, or the strict upper triangle of a
        # diagonal tile
        nrow, ncol, diag = T[:, 3], T[:, 4], T[:, 5] != 0
        rr = np.arange(TILE)
        tri = (np.clip(ncol[:, None] - rr[None] - 1, 0, None) * (rr[None] < nrow[:, None])).sum(1)
        nvalid = np.where(diag, tri, nrow * ncol).astype(np.int64)
        # compact integral storage: per tile NW rows of S = round_up(#valid lanes, 16) doubles,
        # valid lanes packed in lane order; offsets are local to each kind's buffer
        S = (nvalid + 15) // 16 * 16
        nw = np.asarray(W_PER_LANE, dtype=np.int64)[self.tile_kind] if nt else np.zeros(0, np.int64)
        woff = np.zeros(nt, dtype=np.int64)
        self.w_size = [0, 0, 0]
        for k in range(3):
            sel = slice(base[k], base[k + 1])
            blk = S[sel] * nw[sel]
            woff[sel] = np.concatenate([[0], np.cumsum(blk)[:-1]]) if len(blk) else blk
            self.w_size[k] = int(blk.sum())
        self.tile_woff_host = woff
        self.tile_wstride = S
        # partial slots: row r of tile t <-> atom row0 + r, column c <-> atom col0 + c; each atom's
        # slots in increasing slot order (a stable sort by atom of the slot enumeration)
        tid = np.repeat(np.arange(nt), 16)
        s = np.tile(np.arange(16), nt)
        atom = np.where(s < 8, T[tid, 1] + s, T[tid, 2] + (s - 8))
        inrange = np.where(s < 8, s < T[tid, 3], (s - 8) < T[tid, 4])
        atom, slot = atom[inrange], (tid * 16 + s)[inrange]
        order = np.argsort(atom, kind="stable")
        self.csr_idx_host = slot[order].astype(np.int64)
        self.csr_off_host = np.searchsorted(atom[order], np.arange(self.natoms + 1)).astype(np.int64)
        self._lanes = None

    def _lane_arrays(self):
        """Per-lane (tile * 64 + lane) views, built on first use (adapters, tests, reference-order
        outputs; the device engine does not need them): the lane's pair index in the reference's
        (mol, i < j) order or -1, its validity, and the position of its first integral in its
        kind's compact buffer or -1."""
        if self._lanes is None:
            T = self.tiles_host
            r = np.arange(64) // TILE
            c = np.arange(64) % TILE
            i = T[:, 1:2] + r[None]
            j = T[:, 2:3] + c[None]
            valid = (r[None] < T[:, 3:4]) & (c[None] < T[:, 4:5]) & ((T[:, 5:6] == 0) | (c[None] > r[None]))
            mol = T[:, 0:1]
            a0 = self.mol_atom_off[mol]
            li, lj = i - a0, j - a0
            nat = self.nat[mol]
            ref = self.pair_base[mol] + li * nat - li * (li + 1) // 2 + (lj - li - 1)
            rank = np.cumsum(valid, axis=1) - 1
            self._lanes = (np.where(valid, ref, -1).reshape(-1), valid.reshape(-1),
                           np.where(valid, self.tile_woff_host[:, None] + rank, -1).reshape(-1))
        return self._lanes

    @property
    def pair_ref_host(self):
        return self._lane_arrays()[0]

    @property
    def lane_valid(self):
        return self._lane_arrays()[1]

    @property
    def lane_wpos(self):
        return self._lane_arrays()[2]

    def pair_ref_device(self):
        """Device copy of pair_ref_host (int32), made on first use."""
        if "pair_ref" not in self.d:
            self.d["pair_ref"] = torch.as_tensor(np.ascontiguousarray(self.pair_ref_host), dtype=torch.int32,
                                                 device=self.device)
        return self.d["pair_ref"]

    # ----------------------------------------------------------------- device
    def _to_device(self, device_tiles=False):
        dev = self.device
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.int32, device=dev)
        self.d = {
            "mol_atom_off": i32(self.mol_atom_off), "mol_norb": i32(self.norb), "mol_nocc": i32(self.nocc),
            "mol_mat_off": torch.as_tensor(self.mol_mat_off, dtype=torch.int64, device=dev),
            "mol_tile_off": i32(self.mol_tile_off), "atom_mol": i32(self.atom_mol), "atom_ao": i32(self.atom_ao),
            "atom_Z": i32(self.Z),
        }
        if device_tiles:
            nt = sum(self.ntile)
            self.d["tiles"] = torch.zeros(max(4 * nt, 4), dtype=torch.int32, device=dev)
            self.d["tile_woff"] = torch.zeros(max(nt, 1), dtype=torch.int64, device=dev)
            self.d["csr_off"] = torch.zeros(self.natoms + 1, dtype=torch.int32, device=dev)
            self.d["csr_idx"] = torch.zeros(max(self.nslots, 1), dtype=torch.int32, device=dev)
            w_size = torch.zeros(3, dtype=torch.int64, device=dev)
            scratch = torch.empty(nt + (self.natoms + 1) // 2 + 1, dtype=torch.int64, device=dev)
            ntile = (ctypes.c_int * 3)(*self.ntile)
            nheavy = i32(self.nheavy)
            _native.call_raw("pyseqm_build_tiles", self.nmol, self.natoms, _native.ptr(self.d["mol_atom_off"]),
                             _native.ptr(nheavy), _native.ptr(self.d["mol_tile_off"]), ntile,
                             _native.ptr(self.d["tiles"]), _native.ptr(self.d["tile_woff"]), _native.ptr(w_size),
                             _native.ptr(self.d["csr_off"]), _native.ptr(self.d["csr_idx"]), _native.ptr(scratch),
                             _native.stream_ptr(dev))
            self.w_size = [int(v) for v in w_size.cpu()]  # the one host read of the build
        else:
            self.d["csr_off"] = i32(self.csr_off_host)
            self.d["csr_idx"] = i32(self.csr_idx_host)
            self.d["tile_woff"] = torch.as_tensor(self.tile_woff_host if len(self.tile_woff_host) else np.zeros(1, np.int64),
                                                  dtype=torch.int64, device=dev)
            T = self.tiles_host
            packed = np.zeros((len(T), 4), dtype=np.int64)
            if len(T):
                packed[:, 0] = T[:, 0]
                packed[:, 1] = T[:, 1]
                packed[:, 2] = T[:, 2]
                packed[:, 3] = T[:, 3] | (T[:, 4] << 8) | (T[:, 5] << 16)
            self.d["tiles"] = i32(packed.reshape(-1)) if len(T) else torch.zeros(4, dtype=torch.int32, device=dev)
        L = _native.Layout()
        L.nmol, L.natoms, L.npairs = self.nmol, self.natoms, self.npairs
        for k in range(3):
            L.ntile[k] = self.ntile[k]
        for name in ("mol_atom_off", "mol_norb", "mol_nocc", "mol_mat_off", "mol_tile_off", "atom_mol",
                     "atom_ao", "atom_Z", "tiles", "csr_off", "csr_idx", "tile_woff"):
            setattr(L, name, self.d[name].data_ptr())
        # per-molecule Fock build: its LDS slots and the launch order, heaviest molecule first
        # (bytes of the build: 1184 per heavy-heavy pair, 176 heavy-H, 32 H-H, 240 per atom)
        nh, nH = self.nheavy.astype(np.int64), self.nhydro.astype(np.int64)
        work = 1184 * nh * (nh - 1) // 2 + 176 * nh * nH + 32 * nH * (nH - 1) // 2 + 240 * (nh + nH)
        self.d["mol_order"] = i32(np.argsort(-work, kind="stable") if self.nmol else np.zeros(1, np.int64))
        L.max_mol_atoms = int(self.nat.max()) if self.nmol else 0
        L.max_mol_slots = int((10 * nh + nH).max()) if self.nmol else 0
        L.mol_order = self.d["mol_order"].data_ptr()
        self.L = L

    @classmethod
    def from_counts(cls, nheavy, nhydro, nocc, molsize, device):
        """Batch with the orbital layout of (nheavy, nhydro) atoms per molecule (heavy atoms first,
        as pack/unpack assume, pack.py:64-96) and the given occupations; the element of each heavy
        atom is irrelevant to the layout (carbon stands in)."""
        nh = np.asarray(torch.as_tensor(nheavy).cpu(), dtype=np.int64).reshape(-1)
        nH = np.asarray(torch.as_tensor(nhydro).cpu(), dtype=np.int64).reshape(-1)
        sp = np.zeros((len(nh), int(molsize)), dtype=np.int64)
        for m in range(len(nh)):
            sp[m, :nh[m]] = 6
            sp[m, nh[m]:nh[m] + nH[m]] = 1
        tore = np.asarray(TORE)[sp].sum(1).astype(np.int64)
        b = cls(sp, "AM1", device, charges=tore % 2)
        b.set_nocc(nocc)
        return b

    def set_nocc(self, nocc):
        self.nocc = np.asarray(torch.as_tensor(nocc).cpu(), dtype=np.int64).reshape(-1).copy()
        self.d["mol_nocc"] = torch.as_tensor(self.nocc, dtype=torch.int32, device=self.device)
        self.L.mol_nocc = self.d["mol_nocc"].data_ptr()

    # ------------------------------------------------------------- helpers
    def w_gather_index(self, kind, device):
        """(pair index, (nvalid, NW) positions in the kind's integral buffer) of every valid lane
        of the tiles of one kind, or (None, None) when the kind has no tiles."""
        lo, hi = int(self.tile_base[kind]), int(self.tile_base[kind]) + self.ntile[kind]
        if hi == lo:
            return None, None
        sel = slice(lo * 64, hi * 64)
        v = self.lane_valid[sel]
        pos = self.lane_wpos[sel][v]
        S = np.repeat(self.tile_wstride[lo:hi], 64)[v]
        idx = pos[:, None] + np.arange(W_PER_LANE[kind])[None, :] * S[:, None]
        return (torch.as_tensor(self.pair_ref_host[sel][v], dtype=torch.int64, device=device),
                torch.as_tensor(idx, dtype=torch.int64, device=device))

    @property
    def ntiles(self):
        return sum(self.ntile)

    def padded_orbital_index(self):
        """For each packed orbital (molecule-major): (mol, padded row in the reference 4*molsize layout)."""
        mols, rows = [], []
        for a in range(self.natoms):
            m, l = self.atom_mol[a], self.atom_loc[a]
            k = 4 if self.Z[a] > 1 else 1
            mols.extend([m] * k)
            rows.extend(range(4 * l, 4 * l + k))
        return np.asarray(mols), np.asarray(rows)

    def matrix_index(self):
        """Flat packed index + (mol, padded row, padded col) of every matrix element."""
        if not hasattr(self, "_midx"):
            om, orow = self.padded_orbital_index()
            orb_off = np.concatenate([[0], np.cumsum(self.norb)])
            flat, mm, rr, cc = [], [], [], []
            for m in range(self.nmol):
                n = int(self.norb[m])
                rows = orow[orb_off[m]:orb_off[m] + n]
                ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
                flat.append(self.mol_matoff(m) + (ii * n + jj).ravel())
                mm.append(np.full(n * n, m))
                rr.append(rows[ii.ravel()])
                cc.append(rows[jj.ravel()])
            self._midx = tuple(np.concatenate(x) for x in (flat, mm, rr, cc))
        return self._midx

    def mol_matoff(self, m):
        return int(self.mol_mat_off[m])

    def diag_index(self):
        """Flat packed index of every diagonal element and its molecule."""
        mol = np.repeat(np.arange(self.nmol), self.norb)
        first = np.concatenate([[0], np.cumsum(self.norb)])[:-1]
        k = np.arange(len(mol)) - first[mol]
        return self.mol_mat_off[mol] + k * (self.norb[mol] + 1), mol

    def to_padded(self, X):
        """Packed matrices (flat) -> reference layout (nmol, 4*molsize, 4*molsize) on the device."""
        flat, mm, rr, cc = (torch.as_tensor(a, device=X.device) for a in self.matrix_index())
        out = torch.zeros(self.nmol, 4 * self.molsize, 4 * self.molsize, dtype=X.dtype, device=X.device)
        out[mm, rr, cc] = X[flat]
        return out

    def from_padded(self, Xp):
        flat, mm, rr, cc = (torch.as_tensor(a, device=Xp.device) for a in self.matrix_index())
        out = torch.zeros(self.total, dtype=Xp.dtype, device=Xp.device)
        out[flat] = Xp[mm, rr, cc]
        return out

    def atoms_from_padded(self, v):
        """(nmol, molsize, ...) reference layout -> per-atom (natoms, ...); (natoms, ...) passes through."""
        v = torch.as_tensor(v)
        if v.dim() >= 2 and tuple(v.shape[:2]) == (self.nmol, self.molsize) and not (
                v.dim() == 2 and v.shape[0] == self.natoms):
            return v.reshape((-1,) + tuple(v.shape[2:]))[torch.as_tensor(self.real_flat, device=v.device)]
        return v

    def atom_to_padded(self, v, fill=0.0):
        """Per-atom (natoms, ...) -> (nmol, molsize, ...) in the reference's padded atom layout."""
        shape = (self.nmol * self.molsize,) + tuple(v.shape[1:])
        out = torch.full(shape, fill, dtype=v.dtype, device=v.device)
        out[torch.as_tensor(self.real_flat, device=v.device)] = v
        return out.reshape((self.nmol, self.molsize) + tuple(v.shape[1:]))
