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
      double pa[10], pab[4];
      ld_tri10(Pm + (size_t)ai * n + ai, n, pa);
#pragma unroll
      for (int k = 0; k < 10; ++k)
        if (kTriI[k] != kTriJ[k]) pa[k] *= 2.0;
      const double pb = Pm[(size_t)aj * n + aj];
#pragma unroll
      for (int a = 0; a < 4; ++a) pab[a] = Pm[(size_t)(ai + a) * n + aj];
      double K[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        const int mu = kTriI[k], nu = kTriJ[k];
        const double v = wt[(size_t)k * S];
        colv[0] += v * pa[k];
        rowv[k] = v * pb;
        K[mu] += v * pab[nu];
        if (mu != nu) K[nu] += v * pab[mu];
      }
      double f[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double h = Hm[(size_t)(ai + a) * n + aj];
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
  const double es = wave_sum(e);
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
  const double es = wave_sum(e);
  if (lane == 0) etile[tile] = es;
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
    if (lane == 0) {
      double e = 0.0;
#pragma unroll
      for (int k = 0; k < 10; ++k) e += (kTriI[k] == kTriJ[k] ? 0.5 : 1.0) * pd[k] * (hd[k] + acc[k]);
      eatom[a] = e;
    }
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

}  // namespace

namespace {

int launch_fock(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx, const double* w_xh,
                const double* w_hh, const double* P, const double* H, double* F, double* part, double* etile,
                double* eatom, double* eelec, const unsigned char* active, hipStream_t s) {
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
