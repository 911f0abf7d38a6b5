// Device builder of the packed batch's pair tiles (the pair enumeration of the reference Parser,
// seqm/basics.py:219-403, grouped into 8 x 8 atom tiles): from per-molecule atom offsets and
// heavy-atom counts, writes every tile record, the compact integral offset of every tile and each
// atom's list of tile partial slots, with one thread per tile or per atom.  The host supplies only
// O(nmol) prefix sums (the tile range of each molecule per kind, closed-form counts); the records
// are the same, in the same order, as the host enumeration (pyseqm_amd/layout.py).
//
// Tiles of one molecule and kind cover a region of its local atoms: heavy-heavy rows and columns
// [0, nh) (upper triangle of 8 x 8 tiles, row-major), heavy-H rows [0, nh) x columns [nh, nat)
// (all tiles, row-major), H-H rows and columns [nh, nat) (upper triangle).  A diagonal tile
// holding one atom has no pair and is left out (it can only be a triangle's last tile).
#include "common.h"

using namespace pyseqm;

namespace {

constexpr int kW[3] = {100, 10, 1};  // integrals per pair: heavy-heavy, heavy-H, H-H

__device__ inline int cdiv8(int x) { return (x + 7) >> 3; }

// region of kind `kind` in a molecule with nh heavy atoms and nH hydrogens (local atom indices)
__device__ inline void region_of(int kind, int nh, int nH, int& rs, int& nr, int& cs, int& nc) {
  if (kind == 0) {
    rs = 0, nr = nh, cs = 0, nc = nh;
  } else if (kind == 1) {
    rs = 0, nr = nh, cs = nh, nc = nH;
  } else {
    rs = nh, nr = nH, cs = nh, nc = nH;
  }
}

// first tile index of row bi in a row-major upper triangle of nb x nb tiles
__device__ inline int tri_row_start(int bi, int nb) { return bi * nb - bi * (bi - 1) / 2; }

// index of the molecule whose range [off[m], off[m + 1]) holds x (off non-decreasing, nmol + 1 entries)
__device__ inline int find_mol(const int* off, int nmol, int x) {
  int lo = 0, hi = nmol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void k_layout_tiles(int nmol, const int* __restrict__ mol_atom_off, const int* __restrict__ mol_nheavy,
                               const int* __restrict__ mol_tile_off, int ntot, int* __restrict__ tiles,
                               long long* __restrict__ wsize) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntot) return;
  const int nm1 = nmol + 1;
  int kind = 0;
  while (kind < 2 && t >= mol_tile_off[kind * nm1 + nmol]) ++kind;
  const int* off = mol_tile_off + kind * nm1;
  const int m = find_mol(off, nmol, t);
  const int q = t - off[m];
  const int a0 = mol_atom_off[m], nh = mol_nheavy[m], nH = mol_atom_off[m + 1] - a0 - nh;
  int rs, nr, cs, nc;
  region_of(kind, nh, nH, rs, nr, cs, nc);
  const bool tri = kind != 1;
  const int nbc = cdiv8(nc);
  int bi, bj;
  if (tri) {
    bi = 0;
    while (bi + 1 < nbc && tri_row_start(bi + 1, nbc) <= q) ++bi;
    bj = bi + (q - tri_row_start(bi, nbc));
  } else {
    bi = q / nbc;
    bj = q % nbc;
  }
  const int nrow = min(8, nr - 8 * bi), ncol = min(8, nc - 8 * bj);
  const int diag = tri && bi == bj;
  int4 rec;
  rec.x = m;
  rec.y = a0 + rs + 8 * bi;
  rec.z = a0 + cs + 8 * bj;
  rec.w = nrow | (ncol << 8) | (diag << 16);
  reinterpret_cast<int4*>(tiles)[t] = rec;
  const int nvalid = diag ? nrow * (nrow - 1) / 2 : nrow * ncol;
  wsize[t] = (long long)kW[kind] * ((nvalid + 15) & ~15);
}

// Exclusive prefix sums of n values in one workgroup (chunks of blockDim.x with a running carry);
// the total goes to *total.  Used for the integral offsets of each kind's tiles and the atoms'
// slot-list offsets.
template <typename T>
__global__ __launch_bounds__(1024) void k_scan(const T* __restrict__ in, T* __restrict__ out, long long n,
                                              T* __restrict__ total) {
  __shared__ T buf[1024];
  T carry = 0;
  for (long long base = 0; base < n; base += blockDim.x) {
    const long long i = base + threadIdx.x;
    const T v = i < n ? in[i] : T(0);
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {  // Hillis-Steele inclusive scan
      const T add = threadIdx.x >= (unsigned)o ? buf[threadIdx.x - o] : T(0);
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < n) out[i] = carry + buf[threadIdx.x] - v;
    const T chunk = buf[blockDim.x - 1];
    __syncthreads();
    carry += chunk;
  }
  if (threadIdx.x == 0) *total = carry;
}

// slot-list length of atom a, or (fill) its slots in increasing slot order
template <bool FILL>
__device__ inline int atom_slots(int la, int nh, int nH, const int* offk /* per kind: first tile of the molecule */,
                                 int* out) {
  int cnt = 0;
  auto put = [&](int tile, int s) {
    if (FILL) out[cnt] = tile * 16 + s;
    ++cnt;
  };
  // the triangle kind (heavy-heavy for heavy atoms, H-H for hydrogens) and the heavy-H rectangle
  const bool heavy = la < nh;
  const int l = heavy ? la : la - nh;
  const int nreg = heavy ? nh : nH;
  const int R = l >> 3, s = l & 7;
  const int nb = cdiv8(nreg);
  const int tkind = heavy ? 0 : 2;
  const bool last_diag_skipped = (nreg & 7) == 1;
  if (heavy) {
    // heavy-heavy first (lowest tile indices), then heavy-H
    for (int bi = 0; bi < R; ++bi) put(offk[tkind] + tri_row_start(bi, nb) + (R - bi), 8 + s);
    if (!(last_diag_skipped && R == nb - 1)) {
      const int t = offk[tkind] + tri_row_start(R, nb);
      put(t, s);
      put(t, 8 + s);
    }
    for (int bj = R + 1; bj < nb; ++bj) put(offk[tkind] + tri_row_start(R, nb) + (bj - R), s);
    const int nbc = cdiv8(nH);
    for (int bj = 0; bj < nbc; ++bj) put(offk[1] + R * nbc + bj, s);
  } else {
    // heavy-H (as a column), then H-H
    const int nbr = cdiv8(nh), nbc = cdiv8(nH);
    for (int bi = 0; bi < nbr; ++bi) put(offk[1] + bi * nbc + R, 8 + s);
    for (int bi = 0; bi < R; ++bi) put(offk[tkind] + tri_row_start(bi, nb) + (R - bi), 8 + s);
    if (!(last_diag_skipped && R == nb - 1)) {
      const int t = offk[tkind] + tri_row_start(R, nb);
      put(t, s);
      put(t, 8 + s);
    }
    for (int bj = R + 1; bj < nb; ++bj) put(offk[tkind] + tri_row_start(R, nb) + (bj - R), s);
  }
  return cnt;
}

template <bool FILL>
__global__ void k_layout_atoms(int nmol, int natoms, const int* __restrict__ mol_atom_off,
                               const int* __restrict__ mol_nheavy, const int* __restrict__ mol_tile_off,
                               int* __restrict__ cnt, const int* __restrict__ csr_off, int* __restrict__ csr_idx) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natoms) return;
  const int m = find_mol(mol_atom_off, nmol, a);
  const int a0 = mol_atom_off[m], nh = mol_nheavy[m], nH = mol_atom_off[m + 1] - a0 - nh;
  const int nm1 = nmol + 1;
  const int offk[3] = {mol_tile_off[m], mol_tile_off[nm1 + m], mol_tile_off[2 * nm1 + m]};
  if (FILL)
    atom_slots<true>(a - a0, nh, nH, offk, csr_idx + csr_off[a]);
  else
    cnt[a] = atom_slots<false>(a - a0, nh, nH, offk, nullptr);
}

}  // namespace

extern "C" int pyseqm_build_tiles(int nmol, int natoms, const int* mol_atom_off, const int* mol_nheavy,
                                  const int* mol_tile_off, const int* ntile, int* tiles, long long* tile_woff,
                                  long long* w_size, int* csr_off, int* csr_idx, long long* scratch, void* stream) {
  if (nmol < 0 || natoms < 0 || !ntile) return PYSEQM_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int nt = ntile[0] + ntile[1] + ntile[2];
  if (nt > 0) {
    k_layout_tiles<<<(nt + 255) / 256, 256, 0, s>>>(nmol, mol_atom_off, mol_nheavy, mol_tile_off, nt, tiles, scratch);
    int base = 0;
    for (int k = 0; k < 3; ++k) {
      k_scan<long long><<<1, 1024, 0, s>>>(scratch + base, tile_woff + base, ntile[k], w_size + k);
      base += ntile[k];
    }
  } else {
    k_scan<long long><<<1, 1024, 0, s>>>(scratch, tile_woff, 0, w_size);
    k_scan<long long><<<1, 1024, 0, s>>>(scratch, tile_woff, 0, w_size + 1);
    k_scan<long long><<<1, 1024, 0, s>>>(scratch, tile_woff, 0, w_size + 2);
  }
  if (natoms > 0) {
    int* cnt = reinterpret_cast<int*>(scratch + nt);  // natoms ints after the tile sizes
    k_layout_atoms<false><<<(natoms + 255) / 256, 256, 0, s>>>(nmol, natoms, mol_atom_off, mol_nheavy, mol_tile_off,
                                                               cnt, nullptr, nullptr);
    k_scan<int><<<1, 1024, 0, s>>>(cnt, csr_off, natoms, csr_off + natoms);
    k_layout_atoms<true><<<(natoms + 255) / 256, 256, 0, s>>>(nmol, natoms, mol_atom_off, mol_nheavy, mol_tile_off,
                                                              nullptr, csr_off, csr_idx);
  }
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}
