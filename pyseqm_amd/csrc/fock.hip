// Fock build F = H + G(P) for the closed-shell NDDO Hamiltonian on gfx950, FP64.
//
// Pass 1 (one 64-lane wave per 8x8 pair tile, one lane per atom pair i<j):
//   * exchange block  K_ij[mu][la] = -1/2 sum_{nu,si} P_ij[nu][si] (mu nu|la si)
//     -> F_ij = H_ij + K_ij written to both triangles,
//   * Coulomb terms  J_i += sum_{la si} P_jj (mu nu|la si),  J_j += sum_{mu nu} P_ii (..),
//     reduced over the tile's rows / columns with wave shuffles into per-tile partials,
//   * this pair's share of 0.5 sum P (H + F).
// Pass 2 (one thread per atom): F_ii = H_ii + one-center G(P_ii) + the atom's partials in a fixed
// order, and the atom's share of the energy.  Pass 3 sums the energy per molecule.
// Follows fock.py:132-347 (one center 187-231, two center 259-347) and energy.py:26-49;
// w is the compact, tile-major integral storage written by pairs.hip.
#include "common.h"

#include <mutex>

using namespace pyseqm;

namespace pyseqm {
void launch_mol_sum(const pyseqm_layout_t* L, const double* tile_v, const double* atom_v,
                    const unsigned char* active, double* out, hipStream_t s);
}

namespace {

enum { XX = 0, XH = 1, HH = 2 };

// Two consecutive doubles at an 8-byte aligned address: one 16-byte access (global_load/store_dwordx4;
// the rows of the 4 x 4 orbital blocks start at any element of a packed block)
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
__device__ inline d2u ld2(const double* p) { return *reinterpret_cast<const d2u*>(p); }
__device__ inline void st2(double* p, double x, double y) { *reinterpret_cast<d2u*>(p) = d2u{x, y}; }

// The 10 packed elements (kTriI, kTriJ order: rows 0..3 of the lower triangle) of the 4 x 4 block at
// p, row stride n, in 6 loads of row segments.
__device__ inline void ld_tri10(const double* p, int n, double (&v)[10]) {
  const d2u r0 = ld2(p), r1 = ld2(p + n), r2a = ld2(p + 2 * n), r2b = ld2(p + 2 * n + 2);
  const d2u r3a = ld2(p + 3 * n), r3b = ld2(p + 3 * n + 2);
  v[0] = r0[0];
  v[1] = r1[0];
  v[2] = r1[1];
  v[3] = r2a[0];
  v[4] = r2a[1];
  v[5] = r2b[0];
  v[6] = r3a[0];
  v[7] = r3a[1];
  v[8] = r3b[0];
  v[9] = r3b[1];
}

// Heavy-H and H-H tiles, one wave per tile (heavy-heavy tiles: k_fock_xx2 below).
template <int KIND>
__global__ __launch_bounds__(64) void k_fock_tiles(pyseqm_layout_t L, int tile_base, const double* __restrict__ w,
                                                    const double* __restrict__ P, const double* __restrict__ H,
                                                    double* __restrict__ F, double* __restrict__ part,
                                                    double* __restrict__ etile,
                                                    const unsigned char* __restrict__ active) {
  static_assert(KIND == XH || KIND == HH, "heavy-heavy tiles run in k_fock_xx2");
  const int tl = xcd_run(blockIdx.x, gridDim.x);
  const int tile = tile_base + tl;
  const int lane = threadIdx.x;
  const Tile t = load_tile(L.tiles, tile);
  if (active && !active[t.mol]) return;  // uniform over the wave
  int i, j;
  const bool valid = lane_pair(t, lane, i, j);
  constexpr int NR = KIND == HH ? 1 : 10;
  constexpr int NC = 1;
  double rowv[NR], colv[NC];
#pragma unroll
  for (int k = 0; k < NR; ++k) rowv[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NC; ++k) colv[k] = 0.0;
  double e = 0.0;
  if (valid) {
    const long long mo = L.mol_mat_off[t.mol];
    const int n = L.mol_norb[t.mol];
    const int ai = L.atom_ao[i], aj = L.atom_ao[j];
    const double* Pm = P + mo;
    const double* Hm = H + mo;
    double* Fm = F + mo;
    if constexpr (KIND == XH) {
      const int S = tile_wstride(t);
      const double* wt = w + L.tile_woff[tile] + lane_rank(t, lane);
      // every load of the tile is issued before any is used (the scheduler otherwise sinks the
      // diagonal-block loads below the F stores and waits out each one in turn)
      double pa[10], pab[4], hab[4], wv[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) wv[k] = wt[(size_t)k * S];
      ld_tri10(Pm + (size_t)ai * n + ai, n, pa);
      const double pb = Pm[(size_t)aj * n + aj];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        pab[a] = Pm[(size_t)(ai + a) * n + aj];
        hab[a] = Hm[(size_t)(ai + a) * n + aj];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 10; ++k)
        if (kTriI[k] != kTriJ[k]) pa[k] *= 2.0;
      double K[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        const int mu = kTriI[k], nu = kTriJ[k];
        const double v = wv[k];
        colv[0] += v * pa[k];
        rowv[k] = v * pb;
        K[mu] += v * pab[nu];
        if (mu != nu) K[nu] += v * pab[mu];
      }
      double f[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double h = hab[a];
        f[a] = h - 0.5 * K[a];
        Fm[(size_t)(ai + a) * n + aj] = f[a];
        e += pab[a] * (h + f[a]);
      }
      st2(Fm + (size_t)aj * n + ai, f[0], f[1]);
      st2(Fm + (size_t)aj * n + ai + 2, f[2], f[3]);
    } else {
      const double v = w[L.tile_woff[tile] + lane_rank(t, lane)];
      const double pij = Pm[(size_t)ai * n + aj];
      rowv[0] = v * Pm[(size_t)aj * n + aj];
      colv[0] = v * Pm[(size_t)ai * n + ai];
      const double h = Hm[(size_t)ai * n + aj];
      const double f = h - 0.5 * v * pij;
      Fm[(size_t)ai * n + aj] = f;
      Fm[(size_t)aj * n + ai] = f;
      e = pij * (h + f);
    }
  }
  write_partials_c(part, L, tile, lane, rowv, colv);
  const double es = wave64(e);  // every lane active here
  if (lane == 0) etile[tile] = es;
}

// The integral contraction of k_fock_xx2 for the (mu nu) pairs k = K0 .. K0+4 (static indices).
template <int K0>
__device__ inline void xx_half(const double* __restrict__ wt, int S, const double* Pm, int n, int ai, int aj,
                               const double (&pab)[4][4], double (&rowv)[5], double (&colv)[10], double (&K)[4][4]) {
  double pa[5], pb[10];
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) {
    const int k = K0 + kk;
    pa[kk] = (kTriI[k] == kTriJ[k] ? 1.0 : 2.0) * Pm[(size_t)(ai + kTriI[k]) * n + ai + kTriJ[k]];
  }
  ld_tri10(Pm + (size_t)aj * n + aj, n, pb);
#pragma unroll
  for (int l = 0; l < 10; ++l)
    if (kTriI[l] != kTriJ[l]) pb[l] *= 2.0;
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) {
    const int k = K0 + kk;
    const int mu = kTriI[k], nu = kTriJ[k];
#pragma unroll
    for (int l = 0; l < 10; ++l) {
      const int la = kTriI[l], si = kTriJ[l];
      const double v = wt[(size_t)(k * 10 + l) * S];
      colv[l] += v * pa[kk];
      rowv[kk] += v * pb[l];
      K[mu][la] += v * pab[nu][si];
      if (mu != nu) K[nu][la] += v * pab[mu][si];
      if (la != si) K[mu][si] += v * pab[nu][la];
      if (mu != nu && la != si) K[nu][si] += v * pab[mu][la];
    }
  }
}

// Heavy-heavy tiles with two waves per tile: wave q contracts the integrals (kl) with k in its half
// of the 10 (mu nu) pairs, so each wave streams 50 of the 100 integral columns; wave 1 hands its
// partial Coulomb column sums, exchange block and row sums to wave 0 through LDS, and wave 0 writes
// F, the tile partials and the energy.  Per pair: exchange block K from the 4 x 4 off-diagonal P
// block, Coulomb row / column sums from the packed diagonal blocks of P (off-diagonal pairs doubled).
__global__ __launch_bounds__(128) void k_fock_xx2(pyseqm_layout_t L, const double* __restrict__ w,
                                                  const double* __restrict__ P, const double* __restrict__ H,
                                                  double* __restrict__ F, double* __restrict__ part,
                                                  double* __restrict__ etile,
                                                  const unsigned char* __restrict__ active) {
  __shared__ double xch[31][64];
  const int tile = xcd_run(blockIdx.x, gridDim.x);
  const int q = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const Tile t = load_tile(L.tiles, tile);
  if (active && !active[t.mol]) return;  // uniform over the block
  int i, j;
  const bool valid = lane_pair(t, lane, i, j);
  double rowv[5], colv[10], K[4][4];
#pragma unroll
  for (int k = 0; k < 5; ++k) rowv[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 10; ++k) colv[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) K[a][b] = 0.0;
  long long mo = 0;
  int n = 0, ai = 0, aj = 0;
  double pab[4][4];
  if (valid) {
    mo = L.mol_mat_off[t.mol];
    n = L.mol_norb[t.mol];
    ai = L.atom_ao[i];
    aj = L.atom_ao[j];
    const double* Pm = P + mo;
    const int S = tile_wstride(t);
    const double* wt = w + L.tile_woff[tile] + lane_rank(t, lane);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const d2u u0 = ld2(Pm + (size_t)(ai + a) * n + aj), u1 = ld2(Pm + (size_t)(ai + a) * n + aj + 2);
      pab[a][0] = u0[0];
      pab[a][1] = u0[1];
      pab[a][2] = u1[0];
      pab[a][3] = u1[1];
    }
    if (q == 0)
      xx_half<0>(wt, S, Pm, n, ai, aj, pab, rowv, colv, K);
    else
      xx_half<5>(wt, S, Pm, n, ai, aj, pab, rowv, colv, K);
  }
  if (q == 1) {
#pragma unroll
    for (int l = 0; l < 10; ++l) xch[l][lane] = colv[l];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) xch[10 + 4 * a + b][lane] = K[a][b];
#pragma unroll
    for (int kk = 0; kk < 5; ++kk) xch[26 + kk][lane] = rowv[kk];
  }
  __syncthreads();
  if (q == 1) return;
  double row10[10];
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) {
    row10[kk] = rowv[kk];
    row10[5 + kk] = xch[26 + kk][lane];
  }
#pragma unroll
  for (int l = 0; l < 10; ++l) colv[l] += xch[l][lane];
  double e = 0.0;
  if (valid) {
    const double* Hm = H + mo;
    double* Fm = F + mo;
    double f[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const d2u h0 = ld2(Hm + (size_t)(ai + a) * n + aj), h1 = ld2(Hm + (size_t)(ai + a) * n + aj + 2);
      const double h[4] = {h0[0], h0[1], h1[0], h1[1]};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const double kt = K[a][b] + xch[10 + 4 * a + b][lane];
        f[a][b] = h[b] - 0.5 * kt;
        e += pab[a][b] * (h[b] + f[a][b]);
      }
      st2(Fm + (size_t)(ai + a) * n + aj, f[a][0], f[a][1]);
      st2(Fm + (size_t)(ai + a) * n + aj + 2, f[a][2], f[a][3]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      st2(Fm + (size_t)(aj + b) * n + ai, f[0][b], f[1][b]);
      st2(Fm + (size_t)(aj + b) * n + ai + 2, f[2][b], f[3][b]);
    }
  }
  write_partials_c(part, L, tile, lane, row10, colv);
  const double es = wave64(e);  // every lane active here
  if (lane == 0) etile[tile] = es;
}

// One-center G(P_ii) of a heavy atom (fock.py:196-224) added to acc, which holds H_ii plus the
// atom's Coulomb sums (packed lower triangle, kTriI/kTriJ order); returns the atom's energy share
// 1/2 sum P_ii (H_ii + F_ii) over both triangles of the block.
__device__ inline double onecenter_heavy(const double* p, const double (&pd)[10], const double (&hd)[10],
                                         double (&acc)[10]) {
  const double gss = p[PAR_GSS], gsp = p[PAR_GSP], gpp = p[PAR_GPP], gp2 = p[PAR_GP2], hsp = p[PAR_HSP];
  const double pss = pd[0], ppt = pd[2] + pd[5] + pd[9];
  const double f1 = gsp - 0.5 * hsp, f2 = 1.5 * hsp - 0.5 * gsp;
  const double fd = 1.25 * gp2 - 0.25 * gpp, fo = 0.75 * gpp - 1.25 * gp2;
  acc[0] += 0.5 * pss * gss + ppt * f1;
  acc[2] += pss * f1 + 0.5 * pd[2] * gpp + (ppt - pd[2]) * fd;
  acc[5] += pss * f1 + 0.5 * pd[5] * gpp + (ppt - pd[5]) * fd;
  acc[9] += pss * f1 + 0.5 * pd[9] * gpp + (ppt - pd[9]) * fd;
  acc[1] += pd[1] * f2;
  acc[3] += pd[3] * f2;
  acc[6] += pd[6] * f2;
  acc[4] += pd[4] * fo;
  acc[7] += pd[7] * fo;
  acc[8] += pd[8] * fo;
  double e = 0.0;
#pragma unroll
  for (int k = 0; k < 10; ++k) e += (kTriI[k] == kTriJ[k] ? 0.5 : 1.0) * pd[k] * (hd[k] + acc[k]);
  return e;
}

// Diagonal blocks: H_ii + one-center G (fock.py:196-224) + Coulomb partials; energy share.
// One wave per atom: the lanes split the atom's partial slots, then lanes 0..15 write the block.
template <int G>
__global__ __launch_bounds__(256) void k_fock_diag(pyseqm_layout_t L, const double* __restrict__ par,
                                                   const double* __restrict__ P, const double* __restrict__ H,
                                                   const double* __restrict__ part, double* __restrict__ F,
                                                   double* __restrict__ eatom,
                                                   const unsigned char* __restrict__ active) {
  const int a = blockIdx.x * (256 / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (a >= L.natoms) return;  // uniform over the lane group
  const int m = L.atom_mol[a];
  if (active && !active[m]) return;
  const int n = L.mol_norb[m];
  const long long mo = L.mol_mat_off[m];
  const double* Pm = P + mo;
  const double* Hm = H + mo;
  double* Fm = F + mo;
  const int ao = L.atom_ao[a];
  const double* p = par + (size_t)a * PYSEQM_NPAR;
  const int e0 = L.csr_off[a], e1 = L.csr_off[a + 1];
  if (L.atom_Z[a] > 1) {
    double acc[10], pd[10], hd[10];
    gather_slots_c<G>(part, L, e0, e1, lane, acc);
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      pd[k] = Pm[(size_t)(ao + kTriI[k]) * n + ao + kTriJ[k]];
      hd[k] = Hm[(size_t)(ao + kTriI[k]) * n + ao + kTriJ[k]];
      acc[k] += hd[k];
    }
    const double e = onecenter_heavy(p, pd, hd, acc);
    if (lane == 0) eatom[a] = e;
    if (lane < 16) {
      const int mu = lane >> 2, nu = lane & 3;
      double v = acc[0];
#pragma unroll
      for (int k = 1; k < 10; ++k)
        if (k == tri_index(mu, nu)) v = acc[k];
      Fm[(size_t)(ao + mu) * n + ao + nu] = v;
    }
  } else {
    double acc[1];
    gather_slots_c<G>(part, L, e0, e1, lane, acc);
    if (lane == 0) {
      const double pd = Pm[(size_t)ao * n + ao], hd = Hm[(size_t)ao * n + ao];
      const double f = acc[0] + hd + 0.5 * pd * p[PAR_GSS];
      Fm[(size_t)ao * n + ao] = f;
      eatom[a] = 0.5 * pd * (hd + f);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Per-molecule Fock build (molecules of at most PYSEQM_FOCK_MOL_ATOMS atoms): one workgroup of NW
// waves per molecule does the pair tiles, the Coulomb sums and the diagonal blocks in one launch,
// with the sums in LDS instead of tile partials in global memory.
//
// Work items of a molecule, dealt to the waves in a fixed round robin (item k -> wave k mod NW):
// every heavy-heavy tile twice (exchange block and F; Coulomb sums), then groups of two heavy-H
// and of four H-H tiles whose loads are all issued before any is used.  The two passes over a
// heavy-heavy tile stream the same integrals on neighbouring waves, so the second pass reads them
// from L2.  Each wave adds its Coulomb sums into its own per-atom LDS slots (10 per heavy atom, 1
// per hydrogen; heavy atoms first), in its fixed item order, so the sums do not depend on timing;
// after a barrier one thread per atom adds the NW slots in wave order, the one-center terms and H,
// and writes the diagonal block.
//
// Where it pays: a molecule's items run one after another on its NW waves, so the molecule with
// the most items sets a floor under the launch, while the tile-parallel kernels spread every tile
// over the whole chip but pay five launches and the partial round trip through global memory.
// On batches of 1024 molecules the per-molecule kernel is faster up to ~24 atoms per molecule
// (16 benzenes 10 us vs 23 us; (H2O)_8 13 vs 26 us) and slower from ~30 atoms on (alkanes of 2-30
// carbons 171-181 us vs 157 us; (H2O)_40 108-115 vs 72 us), hence the PYSEQM_FOCK_MOL_ATOMS rule.

// (the DPP and permlane-swap sums row8 / col8 / wave64 are in common.h)

// Element k (0..9) of a per-lane 10-vector, k varying over lanes, without dynamic register indexing.
__device__ inline double pick10(const double (&v)[10], int k) {
  double x = v[0];
#pragma unroll
  for (int t = 1; t < 10; ++t) x = k == t ? v[t] : x;
  return x;
}

// Local LDS slot of atom la (index inside the molecule; nh heavy atoms first) and its first orbital
// (4 per heavy atom, then 1 per hydrogen).
__device__ inline int mslot(int la, int nh) { return la < nh ? 10 * la : 9 * nh + la; }
__device__ inline int mao(int la, int nh) { return la < nh ? 4 * la : 3 * nh + la; }

// A pair tile in local atom indices, with the offset of its integrals.
struct MTile {
  int row0, col0, nrow, ncol, diag;
  long long woff;
};
// A tile record read from LDS, made wave-uniform (scalar registers) so that the integral offsets
// and strides are scalar and the loads take a scalar base plus the lane's offset.
__device__ inline MTile uniform_tile(const MTile* tl, int u) {
  const MTile x = tl[u];
  MTile t;
  t.row0 = __builtin_amdgcn_readfirstlane(x.row0);
  t.col0 = __builtin_amdgcn_readfirstlane(x.col0);
  t.nrow = __builtin_amdgcn_readfirstlane(x.nrow);
  t.ncol = __builtin_amdgcn_readfirstlane(x.ncol);
  t.diag = __builtin_amdgcn_readfirstlane(x.diag);
  const unsigned long long wo = (unsigned long long)x.woff;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)wo);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(wo >> 32));
  t.woff = (long long)(((unsigned long long)hi << 32) | lo);
  return t;
}
__device__ inline bool mlane_pair(const MTile& t, int lane, int& i, int& j) {
  const int r = lane >> 3, c = lane & 7;
  i = t.row0 + r;
  j = t.col0 + c;
  return r < t.nrow && c < t.ncol && (!t.diag || c > r);
}
__device__ inline int mlane_rank(const MTile& t, int lane) {
  const int r = lane >> 3, c = lane & 7;
  return t.diag ? r * t.nrow - r * (r + 1) / 2 + (c - r - 1) : r * t.ncol + c;
}
__device__ inline int mwstride(const MTile& t) {
  const int nv = t.diag ? t.nrow * (t.nrow - 1) / 2 : t.nrow * t.ncol;
  return (nv + 15) & ~15;
}

// The 10 integrals (k la si), la si = 0..9, of integral row k of a heavy-heavy pair: wb is the
// tile's (wave-uniform) integral base, S its row stride, rk the lane's rank in the tile; the row
// address is scalar and the lane adds its offset.
__device__ inline void load_row10(const double* __restrict__ wb, int S, unsigned rk, int k, double (&v)[10]) {
#pragma unroll
  for (int l = 0; l < 10; ++l) v[l] = __builtin_nontemporal_load(wb + (size_t)(k * 10 + l) * S + rk);
}

// Heavy-heavy tile, exchange pass: F_ij = H_ij - 1/2 K_ij in both triangles, energy share.
__device__ inline void mol_xx_exchange(const MTile& t, const double* __restrict__ w, const double* Pm,
                                       const double* Hm, double* Fm, int n, int lane, double& e) {
  int i, j;
  if (!mlane_pair(t, lane, i, j)) return;
  const int S = mwstride(t);
  const double* wb = w + t.woff;
  const unsigned rk = (unsigned)mlane_rank(t, lane);
  const int ai = 4 * i, aj = 4 * j;
  // K starts at -2 H, so that F = -K/2 at the end (H is read with P, up front, and not kept)
  double pab[4][4], K[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const d2u u0 = ld2(Pm + (size_t)(ai + a) * n + aj), u1 = ld2(Pm + (size_t)(ai + a) * n + aj + 2);
    const d2u h0 = ld2(Hm + (size_t)(ai + a) * n + aj), h1 = ld2(Hm + (size_t)(ai + a) * n + aj + 2);
    pab[a][0] = u0[0];
    pab[a][1] = u0[1];
    pab[a][2] = u1[0];
    pab[a][3] = u1[1];
    const double h[4] = {h0[0], h0[1], h1[0], h1[1]};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      e += pab[a][b] * h[b];
      K[a][b] = -2.0 * h[b];
    }
  }
  // integral rows k = (mu nu), 10 integrals (la si) each, loaded one row ahead of their use
  double cur[10], nxt[10];
  load_row10(wb, S, rk, 0, cur);
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const int mu = kTriI[k], nu = kTriJ[k];
    if (k < 9) load_row10(wb, S, rk, k + 1, nxt);
#pragma unroll
    for (int l = 0; l < 10; ++l) {
      const int la = kTriI[l], si = kTriJ[l];
      const double v = cur[l];
      K[mu][la] += v * pab[nu][si];
      if (mu != nu) K[nu][la] += v * pab[mu][si];
      if (la != si) K[mu][si] += v * pab[nu][la];
      if (mu != nu && la != si) K[nu][si] += v * pab[mu][la];
    }
#pragma unroll
    for (int l = 0; l < 10; ++l) cur[l] = nxt[l];
  }
  double f[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      f[a][b] = -0.5 * K[a][b];
      e += pab[a][b] * f[a][b];
    }
    st2(Fm + (size_t)(ai + a) * n + aj, f[a][0], f[a][1]);
    st2(Fm + (size_t)(ai + a) * n + aj + 2, f[a][2], f[a][3]);
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    st2(Fm + (size_t)(aj + b) * n + ai, f[0][b], f[1][b]);
    st2(Fm + (size_t)(aj + b) * n + ai + 2, f[2][b], f[3][b]);
  }
}

// Heavy-heavy tile, Coulomb pass: atom i += sum_{la si} P_jj (mu nu|la si), atom j += sum_{mu nu}
// P_ii (mu nu|la si), off-diagonal density elements doubled (fock.py:259-290).
__device__ inline void mol_xx_coulomb(const MTile& t, const double* __restrict__ w, const double* pdg,
                                      double* acc, int lane) {
  int i, j;
  const bool valid = mlane_pair(t, lane, i, j);
  double rowv[10], colv[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) rowv[k] = colv[k] = 0.0;
  if (valid) {
    const int S = mwstride(t);
    const double* wb = w + t.woff;
    const unsigned rk = (unsigned)mlane_rank(t, lane);
    const double* pi = pdg + 10 * i;  // P_ii read from LDS one element per integral row
    const double* pj = pdg + 10 * j;
    double pb[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) pb[k] = (kTriI[k] == kTriJ[k] ? 1.0 : 2.0) * pj[k];
    double cur[10], nxt[10];
    load_row10(wb, S, rk, 0, cur);
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      if (k < 9) load_row10(wb, S, rk, k + 1, nxt);
      const double pa = (kTriI[k] == kTriJ[k] ? 1.0 : 2.0) * pi[k];
#pragma unroll
      for (int l = 0; l < 10; ++l) {
        colv[l] += cur[l] * pa;
        rowv[k] += cur[l] * pb[l];
      }
#pragma unroll
      for (int l = 0; l < 10; ++l) cur[l] = nxt[l];
    }
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    rowv[k] = row8(rowv[k]);
    colv[k] = col8(colv[k]);
  }
  const int r = lane >> 3, c = lane & 7;
  if (r < t.nrow) {  // lane (r, c) adds elements c and c + 8 of row atom r
    double* d = acc + 10 * (t.row0 + r);
    d[c] += pick10(rowv, c);
    if (c < 2) d[8 + c] += pick10(rowv, 8 + c);
  }
  if (c < t.ncol) {  // lane (r, c) adds elements r and r + 8 of column atom c
    double* d = acc + 10 * (t.col0 + c);
    d[r] += pick10(colv, r);
    if (r < 2) d[8 + r] += pick10(colv, 8 + r);
  }
}

// G heavy-H tiles (rows heavy, columns H) u0 .. u0+G-1 of the local list, those below u1: every
// tile's loads are issued before any of them is used.  Per tile: exchange column, F in both
// triangles, Coulomb sums, energy.
template <int G>
__device__ inline void mol_xh(const MTile* tl, int u0, int u1, const double* __restrict__ w, const double* Pm,
                              const double* Hm, double* Fm, int n, const double* pdg, double* acc, int nh, int lane,
                              double& e) {
  MTile t[G];
  bool valid[G];
  int ai[G], aj[G];
  double wv[G][10], pab[G][4], hab[G][4];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    valid[g] = false;
    t[g].nrow = t[g].ncol = 0;
    if (u0 + g < u1) {
      t[g] = uniform_tile(tl, u0 + g);
      int i, j;
      valid[g] = mlane_pair(t[g], lane, i, j);
      ai[g] = 4 * i;
      aj[g] = 3 * nh + j;
      if (valid[g]) {
        const int S = mwstride(t[g]);
        const double* wb = w + t[g].woff;
        const unsigned rk = (unsigned)mlane_rank(t[g], lane);
#pragma unroll
        for (int k = 0; k < 10; ++k) wv[g][k] = __builtin_nontemporal_load(wb + (size_t)k * S + rk);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          pab[g][a] = Pm[(size_t)(ai[g] + a) * n + aj[g]];
          hab[g][a] = Hm[(size_t)(ai[g] + a) * n + aj[g]];
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (u0 + g >= u1) break;  // uniform over the wave
    double rowv[10], colv = 0.0;
#pragma unroll
    for (int k = 0; k < 10; ++k) rowv[k] = 0.0;
    if (valid[g]) {
      const int i = ai[g] / 4, j = aj[g] - 3 * nh;
      const double* pi = pdg + 10 * i;
      const double pb = pdg[9 * nh + j];
      double K[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        const int mu = kTriI[k], nu = kTriJ[k];
        const double v = wv[g][k];
        colv += v * ((mu == nu ? 1.0 : 2.0) * pi[k]);
        rowv[k] = v * pb;
        K[mu] += v * pab[g][nu];
        if (mu != nu) K[nu] += v * pab[g][mu];
      }
      double f[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        f[a] = hab[g][a] - 0.5 * K[a];
        Fm[(size_t)(ai[g] + a) * n + aj[g]] = f[a];
        e += pab[g][a] * (hab[g][a] + f[a]);
      }
      st2(Fm + (size_t)aj[g] * n + ai[g], f[0], f[1]);
      st2(Fm + (size_t)aj[g] * n + ai[g] + 2, f[2], f[3]);
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) rowv[k] = row8(rowv[k]);
    colv = col8(colv);
    const int r = lane >> 3, c = lane & 7;
    if (r < t[g].nrow) {
      double* d = acc + 10 * (t[g].row0 + r);
      d[c] += pick10(rowv, c);
      if (c < 2) d[8 + c] += pick10(rowv, 8 + c);
    }
    if (r == 0 && c < t[g].ncol) acc[9 * nh + t[g].col0 + c] += colv;
  }
}

// G H-H tiles u0 .. u0+G-1 (those below u1), loads first.
template <int G>
__device__ inline void mol_hh(const MTile* tl, int u0, int u1, const double* __restrict__ w, const double* Pm,
                              const double* Hm, double* Fm, int n, const double* pdg, double* acc, int nh, int lane,
                              double& e) {
  MTile t[G];
  bool valid[G];
  int ai[G], aj[G];
  double v[G], pij[G], h[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    valid[g] = false;
    t[g].nrow = t[g].ncol = 0;
    if (u0 + g < u1) {
      t[g] = uniform_tile(tl, u0 + g);
      int i, j;
      valid[g] = mlane_pair(t[g], lane, i, j);
      ai[g] = 3 * nh + i;
      aj[g] = 3 * nh + j;
      if (valid[g]) {
        v[g] = __builtin_nontemporal_load((w + t[g].woff) + (unsigned)mlane_rank(t[g], lane));
        pij[g] = Pm[(size_t)ai[g] * n + aj[g]];
        h[g] = Hm[(size_t)ai[g] * n + aj[g]];
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (u0 + g >= u1) break;  // uniform over the wave
    double rowv = 0.0, colv = 0.0;
    if (valid[g]) {
      rowv = v[g] * pdg[aj[g] + 6 * nh];  // slot 9 nh + j of H atom j, orbital 3 nh + j
      colv = v[g] * pdg[ai[g] + 6 * nh];
      const double f = h[g] - 0.5 * v[g] * pij[g];
      Fm[(size_t)ai[g] * n + aj[g]] = f;
      Fm[(size_t)aj[g] * n + ai[g]] = f;
      e += pij[g] * (h[g] + f);
    }
    rowv = row8(rowv);
    colv = col8(colv);
    const int r = lane >> 3, c = lane & 7;
    if (c == 0 && r < t[g].nrow) acc[9 * nh + t[g].row0 + r] += rowv;
    if (r == 0 && c < t[g].ncol) acc[9 * nh + t[g].col0 + c] += colv;
  }
}

#ifndef PYSEQM_FOCK_MOL_WAVES
#define PYSEQM_FOCK_MOL_WAVES 4  // waves per molecule workgroup
#endif
#ifndef PYSEQM_FOCK_MOL_OCC
#define PYSEQM_FOCK_MOL_OCC 3  // waves per SIMD the register allocation must leave room for
#endif
// Tiles of a molecule of at most PYSEQM_FOCK_MOL_ATOMS atoms: at most B (B + 1) / 2 for the
// B = A / 8 + 1 blocks of 8 atoms that a heavy and a hydrogen run of A atoms in all can span.
constexpr int kMolTiles = (PYSEQM_FOCK_MOL_ATOMS / 8 + 1) * (PYSEQM_FOCK_MOL_ATOMS / 8 + 2) / 2;

template <int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(PYSEQM_FOCK_MOL_OCC))) void k_fock_mol(pyseqm_layout_t L, const double* __restrict__ par,
                                                      const double* __restrict__ w_xx,
                                                      const double* __restrict__ w_xh,
                                                      const double* __restrict__ w_hh, const double* __restrict__ P,
                                                      const double* __restrict__ H, double* __restrict__ F,
                                                      double* __restrict__ eelec,
                                                      const unsigned char* __restrict__ active, int stride) {
  extern __shared__ double lds[];  // [NW][stride] Coulomb slots of each wave, then [stride] diagonal blocks of P
  __shared__ MTile tl[kMolTiles];
  __shared__ double ew[NW];
  const int m = L.mol_order ? L.mol_order[blockIdx.x] : (int)blockIdx.x;
  if (active && !active[m]) return;  // uniform over the block
  const int tid = threadIdx.x, q = tid >> 6, lane = tid & 63;
  const int a0 = L.mol_atom_off[m], na = L.mol_atom_off[m + 1] - a0;
  const int n = L.mol_norb[m];
  const int nh = (n - na) / 3;  // n = 4 nh + (na - nh)
  const long long mo = L.mol_mat_off[m];
  const double* Pm = P + mo;
  const double* Hm = H + mo;
  double* Fm = F + mo;
  double* pdg = lds + NW * stride;
  const int T = L.nmol + 1;
  const int x0 = L.mol_tile_off[m], nxx = L.mol_tile_off[m + 1] - x0;
  const int h0 = L.mol_tile_off[T + m], nxh = L.mol_tile_off[T + m + 1] - h0;
  const int g0 = L.mol_tile_off[2 * T + m], nhh = L.mol_tile_off[2 * T + m + 1] - g0;
  const int ntl = nxx + nxh + nhh;  // <= kMolTiles for at most PYSEQM_FOCK_MOL_ATOMS atoms
  if (na > PYSEQM_FOCK_MOL_ATOMS || ntl > kMolTiles || 9 * nh + na > stride) {
    // a layout whose max_mol_atoms / max_mol_slots understate this molecule: nothing is written
    // beyond the LDS arrays, and Eelec comes back NaN
    if (tid == 0) eelec[m] = __builtin_nan("");
    return;  // uniform over the block
  }
  for (int u = tid; u < ntl; u += 64 * NW) {
    const int tile = u < nxx ? x0 + u : u < nxx + nxh ? h0 + (u - nxx) : g0 + (u - nxx - nxh);
    const Tile g = load_tile(L.tiles, tile);
    MTile t;
    t.row0 = g.row0 - a0;
    t.col0 = g.col0 - a0;
    t.nrow = g.nrow;
    t.ncol = g.ncol;
    t.diag = g.diag;
    t.woff = L.tile_woff[tile];
    tl[u] = t;
  }
  for (int x = tid; x < NW * stride; x += 64 * NW) lds[x] = 0.0;
  if (tid < na) {
    const int ao = mao(tid, nh);
    if (tid < nh) {
      double v[10];
      ld_tri10(Pm + (size_t)ao * n + ao, n, v);
#pragma unroll
      for (int k = 0; k < 10; ++k) pdg[10 * tid + k] = v[k];
    } else {
      pdg[9 * nh + tid] = Pm[(size_t)ao * n + ao];
    }
  }
  __syncthreads();

  // items: two per heavy-heavy tile, then groups of GXH heavy-H and GHH H-H tiles
  constexpr int GXH = 2, GHH = 4;
  const int nixh = (nxh + GXH - 1) / GXH, nihh = (nhh + GHH - 1) / GHH;
  const int items = 2 * nxx + nixh + nihh;
  double* acc = lds + q * stride;
  double e = 0.0;
  for (int it = q; it < items; it += NW) {  // uniform over the wave
    if (it < 2 * nxx) {
      const MTile t = uniform_tile(tl, it >> 1);
      if (it & 1)
        mol_xx_coulomb(t, w_xx, pdg, acc, lane);
      else
        mol_xx_exchange(t, w_xx, Pm, Hm, Fm, n, lane, e);
    } else if (it < 2 * nxx + nixh) {
      const int u0 = nxx + GXH * (it - 2 * nxx);
      mol_xh<GXH>(tl, u0, nxx + nxh, w_xh, Pm, Hm, Fm, n, pdg, acc, nh, lane, e);
    } else {
      const int u0 = nxx + nxh + GHH * (it - 2 * nxx - nixh);
      mol_hh<GHH>(tl, u0, nxx + nxh + nhh, w_hh, Pm, Hm, Fm, n, pdg, acc, nh, lane, e);
    }
  }
  __syncthreads();

  // Diagonal blocks: H_ii + one-center G(P_ii) (fock.py:196-224) + the Coulomb slots of every wave.
  if (tid < na) {
    const int ao = mao(tid, nh);
    const double* p = par + (size_t)(a0 + tid) * PYSEQM_NPAR;
    if (tid < nh) {
      double acc10[10], hd[10];
      const double* pd = pdg + 10 * tid;
      ld_tri10(Hm + (size_t)ao * n + ao, n, hd);
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        double sum = lds[10 * tid + k];
#pragma unroll
        for (int w2 = 1; w2 < NW; ++w2) sum += lds[w2 * stride + 10 * tid + k];
        acc10[k] = sum + hd[k];
      }
      double pdr[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) pdr[k] = pd[k];
      e += onecenter_heavy(p, pdr, hd, acc10);
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        st2(Fm + (size_t)(ao + mu) * n + ao, acc10[tri_index(mu, 0)], acc10[tri_index(mu, 1)]);
        st2(Fm + (size_t)(ao + mu) * n + ao + 2, acc10[tri_index(mu, 2)], acc10[tri_index(mu, 3)]);
      }
    } else {
      const int s = 9 * nh + tid;
      double sum = lds[s];
#pragma unroll
      for (int w2 = 1; w2 < NW; ++w2) sum += lds[w2 * stride + s];
      const double pd = pdg[s], hd = Hm[(size_t)ao * n + ao];
      const double f = sum + hd + 0.5 * pd * p[PAR_GSS];
      Fm[(size_t)ao * n + ao] = f;
      e += 0.5 * pd * (hd + f);
    }
  }
  const double es = wave64(e);
  if (lane == 0) ew[q] = es;
  __syncthreads();
  if (tid == 0) {
    double t = ew[0];
#pragma unroll
    for (int w2 = 1; w2 < NW; ++w2) t += ew[w2];
    eelec[m] = t;
  }
}

}  // namespace

namespace {

int launch_fock(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx, const double* w_xh,
                const double* w_hh, const double* P, const double* H, double* F, double* part, double* etile,
                double* eatom, double* eelec, const unsigned char* active, hipStream_t s) {
  if (L->max_mol_atoms > 0 && L->max_mol_atoms <= PYSEQM_FOCK_MOL_ATOMS) {
    constexpr int NW = PYSEQM_FOCK_MOL_WAVES;
    const int stride = L->max_mol_slots;
    if (stride < 1 || stride > 10 * PYSEQM_FOCK_MOL_ATOMS) return PYSEQM_EINVAL;
    k_fock_mol<NW><<<L->nmol, 64 * NW, (size_t)(NW + 1) * stride * sizeof(double), s>>>(
        *L, atom_par, w_xx, w_xh, w_hh, P, H, F, eelec, active, stride);
    return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
  }
  int base = 0;
  if (L->ntile[0]) k_fock_xx2<<<L->ntile[0], 128, 0, s>>>(*L, w_xx, P, H, F, part, etile, active);
  base += L->ntile[0];
  if (L->ntile[1]) k_fock_tiles<XH><<<L->ntile[1], 64, 0, s>>>(*L, base, w_xh, P, H, F, part, etile, active);
  base += L->ntile[1];
  if (L->ntile[2]) k_fock_tiles<HH><<<L->ntile[2], 64, 0, s>>>(*L, base, w_hh, P, H, F, part, etile, active);
  if (L->natoms) {
    if (lanes_per_atom(L) == 64)
      k_fock_diag<64><<<(L->natoms + 3) / 4, 256, 0, s>>>(*L, atom_par, P, H, part, F, eatom, active);
    else
      k_fock_diag<16><<<(L->natoms + 15) / 16, 256, 0, s>>>(*L, atom_par, P, H, part, F, eatom, active);
  }
  pyseqm::launch_mol_sum(L, etile, eatom, active, eelec, s);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

// Scratch of pyseqm_fock_noalloc: the tile partials, per-tile and per-atom energies of one Fock
// build, in a file-scope device array (one copy per GPU, held until the process ends).  Host code
// never passes the array's name to a kernel (on the host it is a stand-in, not a device address):
// it looks up each device's address once with hipGetSymbolAddress and passes that.
static __device__ __align__(256) double scratch[PYSEQM_FOCK_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

extern "C" int pyseqm_fock(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx,
                           const double* w_xh, const double* w_hh, const double* P, const double* H, double* F,
                           double* part, double* etile, double* eatom, double* eelec,
                           const unsigned char* active, void* stream) {
  if (!L) return PYSEQM_EINVAL;
  return launch_fock(L, atom_par, w_xx, w_xh, w_hh, P, H, F, part, etile, eatom, eelec, active,
                     (hipStream_t)stream);
}

extern "C" long long pyseqm_fock_scratch_doubles(const pyseqm_layout_t* L) {
  if (!L) return -1;
  return 160LL * L->ntile[0] + 88LL * L->ntile[1] + 16LL * L->ntile[2] +
         ((long long)L->ntile[0] + L->ntile[1] + L->ntile[2]) + L->natoms;
}

extern "C" int pyseqm_fock_noalloc(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx,
                                   const double* w_xh, const double* w_hh, const double* P, const double* H,
                                   double* F, double* eelec, const unsigned char* active, void* stream) {
  if (!L) return PYSEQM_EINVAL;
  const long long need = pyseqm_fock_scratch_doubles(L);
  if (need > PYSEQM_FOCK_SCRATCH_DOUBLES) return PYSEQM_ETOOBIG;  // checked before any HIP call
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return PYSEQM_ELAUNCH;
  double* base;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PYSEQM_ELAUNCH;
      g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    base = g_scratch_ptr[dev];
  }
  double* part = base;
  double* etile = part + (160LL * L->ntile[0] + 88LL * L->ntile[1] + 16LL * L->ntile[2]);
  double* eatom = etile + ((long long)L->ntile[0] + L->ntile[1] + L->ntile[2]);
  return launch_fock(L, atom_par, w_xx, w_xh, w_hh, P, H, F, part, etile, eatom, eelec, active,
                     (hipStream_t)stream);
}
