// Internal helpers shared by the gfx950 kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pyseqm_hip.h"

namespace pyseqm {

constexpr double kEV = 27.21;         // eV per hartree used by MOPAC7 / PYSEQM (constants.py:4)
constexpr double kA0 = 0.529167;      // bohr radius in Angstrom (constants.py:9)
constexpr double kOverlapCut = 40.0;  // bohr (constants.py:23)
constexpr int kTile = 8;              // pair tile edge: one 64-lane wave per 8 x 8 tile
constexpr int kPartStride = 10;       // doubles per tile partial slot

// tril pair (mu >= nu) of the 4 sp orbitals, index = mu*(mu+1)/2 + nu
__device__ constexpr int kTriI[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};
__device__ constexpr int kTriJ[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
__host__ __device__ constexpr int tri_index(int a, int b) {
  return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
}

struct Tile {
  int mol, row0, col0, nrow, ncol, diag;
};

__device__ inline Tile load_tile(const int* tiles, int t) {
  const int4 v = reinterpret_cast<const int4*>(tiles)[t];
  Tile d;
  d.mol = v.x;
  d.row0 = v.y;
  d.col0 = v.z;
  d.nrow = v.w & 0xff;
  d.ncol = (v.w >> 8) & 0xff;
  d.diag = (v.w >> 16) & 1;
  return d;
}

// Lane mapping inside a tile: lane = 8*r + c  <->  pair (row0 + r, col0 + c).
__device__ inline bool lane_pair(const Tile& t, int lane, int& i, int& j) {
  const int r = lane >> 3, c = lane & 7;
  i = t.row0 + r;
  j = t.col0 + c;
  return r < t.nrow && c < t.ncol && (!t.diag || c > r);
}

// Compact integral storage: a tile keeps one row of S = round_up(#valid pairs, 16) doubles per
// integral, the valid lanes packed in lane order (row-major over the tile, upper triangle for a
// diagonal tile), starting at the tile's offset tile_woff[tile] in its kind's buffer.
__device__ inline int lane_rank(const Tile& t, int lane) {
  const int r = lane >> 3, c = lane & 7;
  return t.diag ? r * t.nrow - r * (r + 1) / 2 + (c - r - 1) : r * t.ncol + c;
}
__device__ inline int tile_wstride(const Tile& t) {
  const int nv = t.diag ? t.nrow * (t.nrow - 1) / 2 : t.nrow * t.ncol;
  return (nv + 15) & ~15;
}

// One DPP lane permutation of a double (two 32-bit moves).  Every control used with it (quad
// permutations, half-row mirror, 16-lane rotation by 8) names a source lane inside the row for
// every lane, so no lane keeps an old value.
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// v + the value of the lane 16 (32) lanes away, with v_permlane16_swap (v_permlane32_swap): the
// swap leaves the even-row (lower-half) value in one register and the odd-row (upper-half) value
// in the other in both partner lanes, so both add the same two numbers in the same order.
__device__ inline double swap16_sum(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  const double y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
  return x + y;
}
__device__ inline double swap32_sum(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  const double y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
  return x + y;
}
// Sum over the 8 lanes of a tile row (lanes 8r..8r+7): xor 1, xor 2 (quad permutations), then the
// other quad (half-row mirror).  Sum over the 8 lanes of a tile column (c, c+8, ..., c+56): xor 8
// (rotate a 16-lane row by 8), xor 16, xor 32.  Every lane of the row (column) ends with the same
// bits.  All 64 lanes must be active.
__device__ inline double row8(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  return v;
}
__device__ inline double col8(double v) {
  v += dpp_d<0x128>(v);
  v = swap16_sum(v);
  return swap32_sum(v);
}
__device__ inline double wave64(double v) { return col8(row8(v)); }

// Sum over the 8 lanes sharing a row r (lanes 8r..8r+7) / a column c (lanes c, c+8, ...).
__device__ inline double row_sum(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  return v;
}
__device__ inline double col_sum(double v) {
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}
__device__ inline double wave_sum(double v) { return col_sum(row_sum(v)); }

// Write the row and column partial sums of a 10-vector held per lane.
// Slot layout: part[(tile*16 + s)*10 + k], s < 8 rows, s >= 8 columns.
template <int NR, int NC>
__device__ inline void write_partials(double* part, int tile, int lane, const double (&rowv)[NR],
                                      const double (&colv)[NC]) {
  const int r = lane >> 3, c = lane & 7;
  double* prow = part + ((size_t)tile * 16 + r) * kPartStride;
  double* pcol = part + ((size_t)tile * 16 + 8 + c) * kPartStride;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const double rs = row_sum(rowv[k]);
    if (c == 0) prow[k] = rs;
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const double cs = col_sum(colv[k]);
    if (r == 0) pcol[k] = cs;
  }
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).  Give each XCD a
// contiguous run of the work list instead, so that neighbouring tiles (which share row segments
// of P, H and F and the atoms' diagonal blocks) share that XCD's L2.  A bijection of [0, nb).
__device__ inline int xcd_run(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int x = b % 8, i = b / 8;
  return x * per + min(x, rem) + i;
}

// Compact partial layout of the Fock and core-Hamiltonian reductions: a tile keeps 8 row slots
// of NR values, then 8 column slots of NC values, with (NR, NC) = (10, 10) for heavy-heavy,
// (10, 1) for heavy-H and (1, 1) for H-H tiles (an H atom carries one value); the tiles of a kind
// are back to back, kinds in the order XX, XH, HH.  (The force reductions keep 10-wide slots.)
__device__ inline size_t cpart_tile(const pyseqm_layout_t& L, int tile, int& nr, int& nc) {
  const int t0 = L.ntile[0], t1 = L.ntile[1];
  if (tile < t0) {
    nr = 10;
    nc = 10;
    return (size_t)tile * 160;
  }
  if (tile < t0 + t1) {
    nr = 10;
    nc = 1;
    return (size_t)t0 * 160 + (size_t)(tile - t0) * 88;
  }
  nr = 1;
  nc = 1;
  return (size_t)t0 * 160 + (size_t)t1 * 88 + (size_t)(tile - t0 - t1) * 16;
}
__device__ inline size_t cpart_slot(const pyseqm_layout_t& L, int slot) {
  int nr, nc;
  const size_t b = cpart_tile(L, slot >> 4, nr, nc);
  const int s = slot & 15;
  return s < 8 ? b + (size_t)s * nr : b + (size_t)8 * nr + (size_t)(s - 8) * nc;
}
template <int NR, int NC>
__device__ inline void write_partials_c(double* part, const pyseqm_layout_t& L, int tile, int lane,
                                        const double (&rowv)[NR], const double (&colv)[NC]) {
  const int r = lane >> 3, c = lane & 7;
  int nr, nc;
  const size_t b = cpart_tile(L, tile, nr, nc);  // (nr, nc) == (NR, NC) for the tile's kind
  double* prow = part + b + (size_t)r * NR;
  double* pcol = part + b + (size_t)8 * NR + (size_t)c * NC;
  // DPP / permlane sums (no LDS round trips; the values' chains interleave), the same pairings
  // and so the same bits as row_sum / col_sum
  double rs[NR], cs[NC];
#pragma unroll
  for (int k = 0; k < NR; ++k) rs[k] = row8(rowv[k]);
#pragma unroll
  for (int k = 0; k < NC; ++k) cs[k] = col8(colv[k]);
  if (c == 0) {
#pragma unroll
    for (int k = 0; k < NR; ++k) prow[k] = rs[k];
  }
  if (r == 0) {
#pragma unroll
    for (int k = 0; k < NC; ++k) pcol[k] = cs[k];
  }
}

// Sum over the G lanes of an aligned lane group (G = 16 or 64), every lane gets the total.
// G = 16 and 64 run on DPP / permlane swaps (the whole group active, as every caller's exits are
// uniform over its group): the pairings of the xor tree, so the same bits, with no LDS round trip.
template <int G>
__device__ inline double group_sum(double v) {
  if constexpr (G == 16) {
    v += dpp_d<0xB1>(v);   // xor 1
    v += dpp_d<0x4E>(v);   // xor 2
    v += dpp_d<0x141>(v);  // the other quad of the half-row
    return v + dpp_d<0x140>(v);  // the other half-row
  } else if constexpr (G == 64) {
    return col8(row8(v));
  } else {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o);
    return v;
  }
}

// Sum of an atom's tile partial slots csr_idx[e0..e1) by a group of G lanes: group lane l adds
// slots e0 + l, e0 + l + G, ... in order, then a fixed shuffle tree; every lane of the group ends
// with the totals (deterministic).
template <int G, int N>
__device__ inline void gather_slots(const double* part, const int* csr_idx, int e0, int e1, int gl,
                                    double (&acc)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = 0.0;
  for (int e = e0 + gl; e < e1; e += G) {
    const double* src = part + (size_t)csr_idx[e] * kPartStride;
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] += src[k];
  }
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = group_sum<G>(acc[k]);
}

// gather_slots over the compact partial layout (cpart_slot addressing).
template <int G, int N>
__device__ inline void gather_slots_c(const double* part, const pyseqm_layout_t& L, int e0, int e1, int gl,
                                      double (&acc)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = 0.0;
  for (int e = e0 + gl; e < e1; e += G) {
    const double* src = part + cpart_slot(L, L.csr_idx[e]);
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] += src[k];
  }
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = group_sum<G>(acc[k]);
}

// Lanes per atom of the per-atom gathers: 64 when atoms carry many partial slots (large
// molecules), 16 otherwise (average slots per atom = 16 ntiles / natoms, at most 10 per atom
// for small molecules).
inline int lanes_per_atom(const pyseqm_layout_t* L) {
  const long long nt = (long long)L->ntile[0] + L->ntile[1] + L->ntile[2];
  return (L->natoms > 0 && 16 * nt > 32LL * L->natoms) ? 64 : 16;
}

}  // namespace pyseqm
