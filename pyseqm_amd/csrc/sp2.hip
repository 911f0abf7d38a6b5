// SP2 density purification on MFMA (gfx950) and the FP64 MFMA GEMM it is corrected with.
//
// Trace-correcting SP2 (seqm/seqm_functions/SP2.py:9-85): X0 = (hN I - F)/(hN - h1) from the
// Gershgorin bounds, then X <- X^2 if |tr X^2 - N| < |2 tr X - tr X^2 - N| else 2X - X^2, until the
// idempotency error stops falling (k_sp2_decide).  Here
//   * X is kept in FP32 in a zero-padded per-molecule block (leading dimension ld = n rounded up
//     to 32, so K never needs a tail and rows are 16-byte aligned);
//   * X^2 runs on v_mfma_f32_32x32x16_bf16 as the 3-term split hi*hi + hi*lo + lo*hi
//     (X = hi + lo, both bf16) with FP32 accumulation;
//   * X is symmetric: only tiles ti <= tj are computed, the epilogue applies the branch, writes the
//     tile and its mirror, and returns FP64 partial sums of tr(X') and ||X'||_F^2 = tr(X'^2), which
//     is exactly the next iteration's branch test, so no second pass over X is needed;
//   * the branch and stopping decisions are taken on the device (one thread per molecule);
//   * a final McWeeny step P = 2 (3 X^2 - 2 X^3) restores FP64 idempotency: X^2 on the FP64 MFMA
//     GEMM (v_mfma_f64_16x16x4_f64), X^3 = X^2 + X (X^2 - X) with the small correction product on
//     the bf16x3 tile loop.
#include "common.h"

#include <type_traits>

using namespace pyseqm;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {

// per-molecule SP2 state, doubles: trace and squared Frobenius norm of the current X, its
// idempotency error tr X - ||X||_F^2 and those of the two previous iterates, the polynomial of the
// next step (1: X^2, 0: 2X - X^2), stop flag, iteration count
enum { S_TR = 0, S_FRO, S_E0, S_E1, S_E2, S_BRANCH, S_DONE, S_IT, S_NST };

// wave sum on DPP / permlane swaps: the pairings of the xor tree, so the same bits, without LDS
// round trips (every call site has its whole wave active)
__device__ inline double wsum(double v) { return wave64(v); }
__device__ inline double wmax(double v) {
  for (int o = 1; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// ------------------------------------------------------------------ init
// Pass 1: Gershgorin bounds hN = max_i (a_ii + r_i), h1 = min_i (a_ii - r_i) (SP2.py:36-43); one
// wave per row, nb blocks per molecule, per-block max/min written to red[(k*nb+b)*4 + 0/1].
__global__ void k_gersh(pyseqm_layout_t L, const double* __restrict__ F, const int* __restrict__ idx, int nb,
                        double* __restrict__ red) {
  __shared__ double sh[2][4];
  const int k = blockIdx.y;
  const int m = idx[k];
  const int n = L.mol_norb[m];
  const double* Fm = F + L.mol_mat_off[m];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double hi = -1.0e300, lo = 1.0e300;
  for (int r = blockIdx.x * 4 + w; r < n; r += nb * 4) {
    double sabs = 0.0;
    for (int c = lane; c < n; c += 64) sabs += fabs(Fm[(size_t)r * n + c]);
    sabs = wsum(sabs);
    const double d = Fm[(size_t)r * n + r];
    hi = fmax(hi, d + (sabs - fabs(d)));
    lo = fmin(lo, d - (sabs - fabs(d)));
  }
  if (lane == 0) {
    sh[0][w] = hi;
    sh[1][w] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0][0], b = sh[1][0];
    for (int q = 1; q < 4; ++q) {
      a = fmax(a, sh[0][q]);
      b = fmin(b, sh[1][q]);
    }
    red[((size_t)k * nb + blockIdx.x) * 4] = a;
    red[((size_t)k * nb + blockIdx.x) * 4 + 1] = b;
  }
}

__device__ inline void gersh_bounds(const double* red, int k, int nb, double* sh2) {
  if (threadIdx.x == 0) {
    double a = -1.0e300, b = 1.0e300;
    for (int q = 0; q < nb; ++q) {
      a = fmax(a, red[((size_t)k * nb + q) * 4]);
      b = fmin(b, red[((size_t)k * nb + q) * 4 + 1]);
    }
    sh2[0] = a;
    sh2[1] = b;
  }
  __syncthreads();
}

// Pass 2: X0 = (hN I - F) / (hN - h1) in FP32 (padded block), per-block FP64 trace / Frobenius
// partials of the rounded X0 in red[(k*nb+b)*4 + 2/3].
__global__ void k_sp2_x0(pyseqm_layout_t L, const double* __restrict__ F, const int* __restrict__ idx, int nb,
                         const long long* __restrict__ xoff, const int* __restrict__ xld, float* __restrict__ X,
                         double* __restrict__ red) {
  __shared__ double sh2[2];
  __shared__ double sh[2][4];
  const int k = blockIdx.y;
  const int m = idx[k];
  const int n = L.mol_norb[m];
  const int ld = xld[m];
  const double* Fm = F + L.mol_mat_off[m];
  float* Xm = X + xoff[m];
  gersh_bounds(red, k, nb, sh2);
  const double hN = sh2[0], inv = 1.0 / (sh2[0] - sh2[1]);
  double tr = 0.0, fro = 0.0;
  // 32-bit element index (blocks are far below 2^32 elements): cheap division into row, column
  const unsigned n2 = (unsigned)ld * (unsigned)ld;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (unsigned)nb * blockDim.x) {
    const int r = (int)(e / (unsigned)ld), c = (int)(e % (unsigned)ld);
    float v = 0.0f;
    if (r < n && c < n) v = (float)(((r == c ? hN : 0.0) - Fm[(size_t)r * n + c]) * inv);
    Xm[e] = v;
    fro += (double)v * v;
    if (r == c) tr += v;
  }
  tr = wsum(tr);
  fro = wsum(fro);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sh[0][w] = tr;
    sh[1][w] = fro;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    red[((size_t)k * nb + blockIdx.x) * 4 + 2] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    red[((size_t)k * nb + blockIdx.x) * 4 + 3] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// Pass 3: per-molecule SP2 state (fixed-order sums of the partials).
__global__ void k_sp2_state0(const int* __restrict__ idx, int nact, int nb, const int* __restrict__ nocc,
                             const double* __restrict__ red, double* __restrict__ st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nact) return;
  const int m = idx[k];
  double t = 0.0, f = 0.0;
  for (int q = 0; q < nb; ++q) {
    t += red[((size_t)k * nb + q) * 4 + 2];
    f += red[((size_t)k * nb + q) * 4 + 3];
  }
  double* s = st + (size_t)m * S_NST;
  const double N = nocc[m];
  s[S_TR] = t;
  s[S_FRO] = f;
  s[S_E0] = s[S_E1] = s[S_E2] = fabs(t - f);
  s[S_BRANCH] = fabs(f - N) < fabs(2.0 * t - f - N) ? 1.0 : 0.0;
  s[S_DONE] = 0.0;
  s[S_IT] = 0.0;
}

// Zero-padded (nmax x nmax) copy of F with the padded diagonal above the Gershgorin range
// (hN + (1 + 0.005 t) dE, diag.py:172-192) for a batched eigensolver.
__global__ void k_pad_fill(pyseqm_layout_t L, const double* __restrict__ F, const int* __restrict__ idx, int nb,
                           int nmax, const double* __restrict__ red, double* __restrict__ out) {
  __shared__ double sh2[2];
  const int k = blockIdx.y;
  const int m = idx[k];
  const int n = L.mol_norb[m];
  const double* Fm = F + L.mol_mat_off[m];
  gersh_bounds(red, k, nb, sh2);
  const double top = sh2[0], dE = sh2[0] - sh2[1];
  double* O = out + (size_t)k * nmax * nmax;
  const unsigned n2 = (unsigned)nmax * (unsigned)nmax;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (unsigned)nb * blockDim.x) {
    const int r = (int)(e / (unsigned)nmax), c = (int)(e % (unsigned)nmax);
    double v = 0.0;
    if (r < n && c < n)
      v = Fm[(size_t)r * n + c];
    else if (r == c)
      v = top + (1.0 + 0.005 * (r - n + 1)) * dE;
    O[e] = v;
  }
}

// ------------------------------------------------------------- BF16x3 GEMM
constexpr int BT = 128;       // output tile edge
constexpr int BK = 32;        // K slab
constexpr int LDSK = BK + 8;  // padded LDS row (bf16): 80 bytes, 16-byte aligned
constexpr int kSp2SmallTileMaxN = 256;  // largest active block that still uses 64 x 64 SP2 tiles

// upper-triangle tile enumeration t = tj*(tj+1)/2 + ti, ti <= tj
__device__ inline void tri_decode(int t, int& ti, int& tj) {
  tj = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((tj + 1) * (tj + 2) / 2 <= t) ++tj;
  while (tj * (tj + 1) / 2 > t) --tj;
  ti = t - tj * (tj + 1) / 2;
}

// One SP2 iteration Y = X^2 (or 2X - X^2) over the upper-triangle TB x TB output tiles of every
// active molecule: 4 waves in 2 x 2, each TB/2 x TB/2 of 32 x 32 MFMA blocks.  Consecutive entries
// of the molecule-major tile list are given to one XCD (block id mod 8), so the tiles that share a
// molecule's row panels share that XCD's L2.
template <int TB>
__global__ __launch_bounds__(256, TB == 128 ? 3 : 6) void k_sp2_bf16x3(
    const int* __restrict__ idx, int nact, int ntri, const int* __restrict__ norb,
    const long long* __restrict__ xoff, const int* __restrict__ xld, const float* __restrict__ Xin,
    float* __restrict__ Xout, double* __restrict__ st, double* __restrict__ part, int part_stride) {
  // K slab of 16: 48 KiB of LDS at TB = 128 (three blocks per CU), 24 KiB at TB = 64 (six)
  constexpr int BK = 16;
  constexpr int LDSK = BK + 8;             // padded LDS row (bf16), 16-byte aligned, conflict-free
  constexpr int NW = TB / 64;              // 32 x 32 blocks per wave side
  constexpr int NQ = TB * (BK / 4) / 256;  // float4 slab loads per thread and operand
  // double-buffered slabs: [buffer][hi/lo][row][k]; after the main loop the same LDS holds the
  // output tile for the mirrored write (in NP passes of TB / NP rows when the whole tile does not
  // fit) and the per-wave trace / Frobenius partials
  constexpr int SLAB_BYTES = 2 * 2 * 2 * TB * LDSK * 2;
  constexpr int NP = TB * (TB + 1) * 4 + 80 <= SLAB_BYTES ? 1 : 2;
  constexpr int PR = TB / NP;              // tile rows per mirror pass
  __shared__ __attribute__((aligned(16))) __bf16 smem[2 * 2 * 2 * TB * LDSK];
  __bf16(*As)[2][TB][LDSK] = reinterpret_cast<__bf16(*)[2][TB][LDSK]>(smem);
  __bf16(*Bs)[2][TB][LDSK] = reinterpret_cast<__bf16(*)[2][TB][LDSK]>(smem + 2 * 2 * TB * LDSK);
  float* T = reinterpret_cast<float*>(smem);
  double(*red)[4] = reinterpret_cast<double(*)[4]>(T + PR * (TB + 1));
  int* zsh = reinterpret_cast<int*>(T + PR * (TB + 1) + 16);
  static_assert(PR * (TB + 1) * 4 + 8 * 8 + 4 <= SLAB_BYTES, "epilogue buffers exceed the slab LDS");
  const int total = nact * ntri, per = (total + 7) / 8;
  const int w = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (w >= total) return;
  const int tile = w % ntri;
  const int m = idx[w / ntri];
  const double* s = st + (size_t)m * S_NST;
  if (s[S_DONE] != 0.0) return;
  const int n = norb[m];
  const int ld = xld[m];
  int ti, tj;
  tri_decode(tile, ti, tj);
  if (tj * TB >= n) return;
  const bool sq = s[S_BRANCH] != 0.0;
  const float* Xm = Xin + xoff[m];
  float* Ym = Xout + xoff[m];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int i0 = ti * TB, j0 = tj * TB;
  floatx16 acc[NW][NW];
#pragma unroll
  for (int a = 0; a < NW; ++a)
#pragma unroll
    for (int b = 0; b < NW; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;

  // global -> registers (next slab, in flight while the MFMAs of this slab run) -> LDS as bf16 hi/lo
  float4 ra[NQ], rb[NQ];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int f = tid + 256 * q;  // float4 index in the TB x (BK/4) slab
      const int row = f / (BK / 4), kq = f % (BK / 4);
      // rows at or beyond n are clamped to row n - 1: they only feed output rows / columns >= n,
      // which are never written (no branches around the loads)
      ra[q] = *reinterpret_cast<const float4*>(Xm + (size_t)min(i0 + row, n - 1) * ld + k0 + 4 * kq);
      rb[q] = *reinterpret_cast<const float4*>(Xm + (size_t)min(j0 + row, n - 1) * ld + k0 + 4 * kq);
    }
  };
  auto split_store = [&](const float4& v, __bf16 (*hi)[LDSK], __bf16 (*lo)[LDSK], int row, int col) {
    union {
      __bf16 h[4];
      uint2 u;
    } ph, pl;
    const float x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ph.h[u] = (__bf16)x[u];
      pl.h[u] = (__bf16)(x[u] - (float)ph.h[u]);
    }
    *reinterpret_cast<uint2*>(&hi[row][col]) = ph.u;
    *reinterpret_cast<uint2*>(&lo[row][col]) = pl.u;
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int f = tid + 256 * q;
      const int row = f / (BK / 4), kq = f % (BK / 4);
      split_store(ra[q], As[buf][0], As[buf][1], row, 4 * kq);
      split_store(rb[q], Bs[buf][0], Bs[buf][1], row, 4 * kq);
    }
  };
  // one barrier per slab: slab t+1 is staged into the other buffer while slab t is multiplied
  gload(0);
  stage(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < ld; k0 += BK) {
    const bool more = k0 + BK < ld;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      bf16x8 ah[NW], al[NW], bh[NW], bl[NW];
      const int kofs = kk + 8 * (lane >> 5);
#pragma unroll
      for (int a = 0; a < NW; ++a) {
        const int r = wr * (TB / 2) + a * 32 + (lane & 31);
        ah[a] = *reinterpret_cast<const bf16x8*>(&As[cur][0][r][kofs]);
        al[a] = *reinterpret_cast<const bf16x8*>(&As[cur][1][r][kofs]);
        const int c = wc * (TB / 2) + a * 32 + (lane & 31);
        bh[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][0][c][kofs]);
        bl[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][1][c][kofs]);
      }
#pragma unroll
      for (int a = 0; a < NW; ++a)
#pragma unroll
        for (int b = 0; b < NW; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        }
    }
    if (more) stage(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // epilogue: Y = X^2 or 2X - X^2, FP64 partial trace / Frobenius norm; the tile is written
  // directly and kept in LDS for the mirrored (transposed) write.  Pass h takes the waves whose
  // rows fall in tile rows [h PR, (h+1) PR).
  double tr = 0.0, fro = 0.0;
  if (tid == 0) *zsh = 0;
  __syncthreads();
#pragma unroll 1
  for (int h = 0; h < NP; ++h) {
    // the element indices are offset by this zero, read from LDS in every pass, so that they are
    // formed inside the pass loop (hoisting them for both passes would exhaust the registers)
    const int z = *reinterpret_cast<volatile int*>(zsh);
    if ((wr * (TB / 2)) / PR == h + z) {
#pragma unroll
      for (int a = 0; a < NW; ++a)
#pragma unroll
        for (int b = 0; b < NW; ++b) {
          // X of the 2X - X^2 branch for this 32 x 32 block: all 16 loads issue before the first
          // use (clamped addresses keep the masked-off elements in bounds)
          float xv[16];
          if (!sq) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
              const int r = min(z + i0 + wr * (TB / 2) + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), n - 1);
              const int c = min(z + j0 + wc * (TB / 2) + b * 32 + (lane & 31), n - 1);
              xv[v] = Xm[(size_t)r * ld + c];
            }
          }
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int r = z + i0 + wr * (TB / 2) + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            const int c = z + j0 + wc * (TB / 2) + b * 32 + (lane & 31);
            if (r < n && c < n && (ti != tj || r <= c)) {
              const float x2 = acc[a][b][v];
              const float y = sq ? x2 : 2.0f * xv[v] - x2;
              Ym[(size_t)r * ld + c] = y;
              T[(r - i0 - h * PR) * (TB + 1) + (c - j0)] = y;
              const double wgt = (r == c) ? 1.0 : 2.0;
              fro += wgt * (double)y * (double)y;
              if (r == c) tr += y;
            }
          }
        }
    }
    __syncthreads();
    {
      // mirror: Y[j0 + cc][i0 + rr] = tile[rr][cc], consecutive lanes along rr (coalesced rows of
      // Y, conflict-free LDS reads with the TB+1 float row pitch)
      const int rlo = h * PR;
      const int rmax = min(TB, n - i0), cmax = min(TB, n - j0);
      for (int e = tid; e < PR * TB; e += 256) {
        const int cc = e / PR, rr = rlo + e % PR;
        if (rr < rmax && cc < cmax && (ti != tj || rr < cc))
          Ym[(size_t)(j0 + cc) * ld + i0 + rr] = T[(rr - rlo) * (TB + 1) + cc];
      }
    }
    __syncthreads();
  }
  tr = wsum(tr);
  fro = wsum(fro);
  if (lane == 0) {
    red[0][wave] = tr;
    red[1][wave] = fro;
  }
  __syncthreads();
  if (tid == 0) {
    double* p = part + (size_t)m * part_stride + 2 * tile;
    p[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    p[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// One TB x TB output tile of sum_p sgn_p A_p[I] B_p[J]^T on bf16x3 MFMA with FP32 accumulation,
// over npass (1 or 2) products: the rows i0 .. i0+TB-1 of A_p and j0 .. j0+TB-1 of B_p are FP32
// with leading dimension ld, zero-padded to ld columns (rows at or beyond n are read as row n - 1:
// only output rows / columns >= n depend on them, and callers write none of those); pass 1's
// A operand is negated (exact in the hi / lo split).  4 waves in 2 x 2, each TB/2 x TB/2 of 32 x 32
// MFMA blocks; K slabs of 16 staged through double-buffered LDS as bf16 hi / lo (row pitch 24), the
// next slab's global loads in flight while the current slab is multiplied.  smem holds
// 2 x 2 x 2 x TB x 24 bf16 (48 KiB at TB = 128).
template <int TB>
__device__ inline void bf16x3_tile(__bf16* smem, floatx16 (&acc)[TB / 64][TB / 64], const float* A0,
                                   const float* B0, const float* A1, const float* B1, int npass, int n, int ld,
                                   int i0, int j0) {
  constexpr int BK = 16, LDSK = BK + 8, NW = TB / 64, NQ = TB * (BK / 4) / 256;
  __bf16(*As)[2][TB][LDSK] = reinterpret_cast<__bf16(*)[2][TB][LDSK]>(smem);
  __bf16(*Bs)[2][TB][LDSK] = reinterpret_cast<__bf16(*)[2][TB][LDSK]>(smem + 2 * 2 * TB * LDSK);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
  for (int a = 0; a < NW; ++a)
#pragma unroll
    for (int b = 0; b < NW; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;
  float4 ra[NQ], rb[NQ];
  // K step t = 0 .. npass nslab - 1: pass t / nslab, slab t % nslab
  const int nslab = ld / BK;
  auto gload = [&](int t) {
    const int pass = t >= nslab;
    const int k0 = (t - pass * nslab) * BK;
    const float* Am = pass ? A1 : A0;
    const float* Bm = pass ? B1 : B0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int f = tid + 256 * q;
      const int row = f / (BK / 4), kq = f % (BK / 4);
      // rows at or beyond n clamped to row n - 1 (they only feed outputs that are never written)
      ra[q] = *reinterpret_cast<const float4*>(Am + (size_t)min(i0 + row, n - 1) * ld + k0 + 4 * kq);
      rb[q] = *reinterpret_cast<const float4*>(Bm + (size_t)min(j0 + row, n - 1) * ld + k0 + 4 * kq);
    }
  };
  auto split_store = [&](const float4& v, float sgn, __bf16 (*hi)[LDSK], __bf16 (*lo)[LDSK], int row, int col) {
    union {
      __bf16 h[4];
      uint2 u;
    } ph, pl;
    const float x[4] = {sgn * v.x, sgn * v.y, sgn * v.z, sgn * v.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ph.h[u] = (__bf16)x[u];
      pl.h[u] = (__bf16)(x[u] - (float)ph.h[u]);
    }
    *reinterpret_cast<uint2*>(&hi[row][col]) = ph.u;
    *reinterpret_cast<uint2*>(&lo[row][col]) = pl.u;
  };
  auto stage = [&](int buf, int t) {
    const float sgn = t >= nslab ? -1.0f : 1.0f;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int f = tid + 256 * q;
      const int row = f / (BK / 4), kq = f % (BK / 4);
      split_store(ra[q], sgn, As[buf][0], As[buf][1], row, 4 * kq);
      split_store(rb[q], 1.0f, Bs[buf][0], Bs[buf][1], row, 4 * kq);
    }
  };
  gload(0);
  stage(0, 0);
  __syncthreads();
  int cur = 0;
  for (int t = 0; t < npass * nslab; ++t) {
    const bool more = t + 1 < npass * nslab;
    if (more) gload(t + 1);
    {
      bf16x8 ah[NW], al[NW], bh[NW], bl[NW];
      const int kofs = 8 * (lane >> 5);
#pragma unroll
      for (int a = 0; a < NW; ++a) {
        const int r = wr * (TB / 2) + a * 32 + (lane & 31);
        ah[a] = *reinterpret_cast<const bf16x8*>(&As[cur][0][r][kofs]);
        al[a] = *reinterpret_cast<const bf16x8*>(&As[cur][1][r][kofs]);
        const int c = wc * (TB / 2) + a * 32 + (lane & 31);
        bh[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][0][c][kofs]);
        bl[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][1][c][kofs]);
      }
#pragma unroll
      for (int a = 0; a < NW; ++a)
#pragma unroll
        for (int b = 0; b < NW; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        }
    }
    if (more) stage(cur ^ 1, t + 1);
    __syncthreads();
    cur ^= 1;
  }
}

// DIIS commutator C = F P - P F on bf16x3 MFMA for the mixed-precision density step, over the
// upper-triangle TB x TB tiles (I <= J) of every listed block only (TB = 64 while the largest
// block has at most 256 orbitals, else 128).  With F and P symmetric,
// C_IJ = sum_k F[I][k] P[J][k] - sum_k P[I][k] F[J][k]: two K passes into one FP32 accumulator,
// the second with its A operand negated.  Operands are FP32 zero-padded copies of F and P in the
// SP2 X layout (xoff, ld = xld).  The epilogue writes the tile's strict upper part of C (FP64) into
// the packed block (offset poff, leading dimension n) and the tile's max |C| to
// tmax[k * tstride + tile].
template <int TB>
__global__ __launch_bounds__(256, TB == 128 ? 3 : 6) void k_comm_bf16x3(const int* __restrict__ idx, int nact,
                                                                       int ntri,
                                                       const int* __restrict__ norb,
                                                       const long long* __restrict__ xoff,
                                                       const int* __restrict__ xld, const float* __restrict__ Fx,
                                                       const float* __restrict__ Px,
                                                       const long long* __restrict__ poff, double* __restrict__ C,
                                                       double* __restrict__ tmax, int tstride) {
  constexpr int LDSK = 16 + 8, NW = TB / 64;
  __shared__ __attribute__((aligned(16))) __bf16 smem[2 * 2 * 2 * TB * LDSK];
  __shared__ double wmx[4];
  const int total = nact * ntri, per = (total + 7) / 8;
  const int w = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (w >= total) return;
  const int k = w / ntri;
  const int tile = w % ntri;
  const int m = idx[k];
  const int n = norb[m];
  const int ld = xld[m];
  int ti, tj;
  tri_decode(tile, ti, tj);
  if (tj * TB >= n) return;  // not one of the molecule's tiles (its max slot is never read)
  const float* Fm = Fx + xoff[m];
  const float* Pm = Px + xoff[m];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int i0 = ti * TB, j0 = tj * TB;
  floatx16 acc[NW][NW];
  bf16x3_tile<TB>(smem, acc, Fm, Pm, Pm, Fm, 2, n, ld, i0, j0);
  double* Cm = C + poff[m];
  double mx = 0.0;
#pragma unroll
  for (int a = 0; a < NW; ++a)
#pragma unroll
    for (int b = 0; b < NW; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int r = i0 + wr * (TB / 2) + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int c = j0 + wc * (TB / 2) + b * 32 + (lane & 31);
        if (r < n && c < n && r < c) {
          const double y = (double)acc[a][b][v];
          Cm[(size_t)r * n + c] = y;
          mx = fmax(mx, fabs(y));
        }
      }
  mx = wmax(mx);
  if (lane == 0) wmx[wave] = mx;
  __syncthreads();
  if (tid == 0) tmax[(size_t)k * tstride + tile] = fmax(fmax(wmx[0], wmx[1]), fmax(wmx[2], wmx[3]));
}

// Per listed molecule: the max over its tiles' max |C| (fixed order) -> diis_err[m].
__global__ __launch_bounds__(64) void k_comm_final(const int* __restrict__ idx, const int* __restrict__ norb,
                                                   const double* __restrict__ tmax, int tstride, int tb,
                                                   double* __restrict__ diis_err) {
  const int k = blockIdx.x;
  const int m = idx[k];
  const int nt = (norb[m] + tb - 1) / tb;
  const int ntri = nt * (nt + 1) / 2;
  double mx = 0.0;
  for (int t = threadIdx.x; t < ntri; t += 64) mx = fmax(mx, tmax[(size_t)k * tstride + t]);
  mx = wmax(mx);
  if (threadIdx.x == 0) diis_err[m] = mx;
}

// Second half of the McWeeny step P = 2 (3 X^2 - 2 X^3) of the final SP2 matrix X, written as
// P = 2 (Y2 - 2 X D) with Y2 = X^2 (FP64) and D = Y2 - X (X^3 = X Y2 = X^2 + X D): D is of the size
// of X's idempotency error (~1e-6 after SP2), so X D on bf16x3 MFMA (relative error ~1e-6) adds
// ~1e-12 to P and the step keeps FP64 accuracy with one FP64 GEMM instead of two.  X is each
// molecule's FP32 block in Xa or Xb by the parity of its iteration count, D (FP32, both triangles)
// the other buffer; Y2 (FP64, X layout) is read on the upper-triangle tiles.  Upper-triangle
// TB x TB tiles; the epilogue writes P to the packed block (offset poff, leading dimension n) and
// mirrors each tile through LDS in passes of 32 rows as coalesced rows of the lower triangle.
template <int TB>
__global__ __launch_bounds__(256, TB == 128 ? 3 : 6) void k_mcw_bf16x3(
    const int* __restrict__ idx, int nact, int ntri, const int* __restrict__ norb,
    const long long* __restrict__ xoff, const int* __restrict__ xld, const float* __restrict__ Xa,
    const float* __restrict__ Xb, const double* __restrict__ st, const double* __restrict__ Y2,
    const long long* __restrict__ poff, double* __restrict__ P) {
  constexpr int LDSK = 16 + 8, NW = TB / 64;
  constexpr int PR = 32, NP = TB / PR, LDT = TB + 1;  // mirror passes of PR tile rows, FP64 LDS pitch
  constexpr int SLAB_BYTES = 2 * 2 * 2 * TB * LDSK * 2;
  __shared__ __attribute__((aligned(16))) __bf16 smem[2 * 2 * 2 * TB * LDSK];
  const int total = nact * ntri, per = (total + 7) / 8;
  const int w = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (w >= total) return;
  const int tile = w % ntri;
  const int m = idx[w / ntri];
  const int n = norb[m];
  const int ld = xld[m];
  int ti, tj;
  tri_decode(tile, ti, tj);
  if (tj * TB >= n) return;
  const bool odd = (((long long)st[(size_t)m * S_NST + S_IT]) & 1) != 0;
  const float* Xm = (odd ? Xb : Xa) + xoff[m];
  const float* Dm = (odd ? Xa : Xb) + xoff[m];
  const double* Ym = Y2 + xoff[m];
  double* Pm = P + poff[m];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int i0 = ti * TB, j0 = tj * TB;
  floatx16 acc[NW][NW];
  // (X D)_IJ = sum_k X[I][k] D[J][k] (D symmetric)
  bf16x3_tile<TB>(smem, acc, Xm, Dm, Xm, Dm, 1, n, ld, i0, j0);
  double* T = reinterpret_cast<double*>(smem);
  int* zsh = reinterpret_cast<int*>(T + PR * LDT);
  static_assert(PR * LDT * 8 + 4 <= SLAB_BYTES, "mirror buffer exceeds the slab LDS");
  if (tid == 0) *zsh = 0;
  __syncthreads();
  const int rmax = min(TB, n - i0), cmax = min(TB, n - j0);
#pragma unroll 1
  for (int h = 0; h < NP; ++h) {
    // the element indices are offset by this zero, read from LDS in every pass, so that they are
    // formed inside the pass loop (hoisting them for all passes would exhaust the registers)
    const int z = *reinterpret_cast<volatile int*>(zsh);
    // pass h: tile rows [h PR, (h+1) PR) = the 32-row MFMA block a = ha of the waves in row wr = hw;
    // those waves put their X D block into LDS
    const int hw = (h * PR) / (TB / 2), ha = ((h * PR) % (TB / 2)) / 32;
    if (wr == hw + z) {
#pragma unroll
      for (int a = 0; a < NW; ++a) {
        if (a != ha) continue;
#pragma unroll
        for (int b = 0; b < NW; ++b)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int rr = z + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);  // row in the pass
            const int cc = z + wc * (TB / 2) + b * 32 + (lane & 31);
            T[(rr - a * 32) * LDT + cc] = (double)acc[a][b][v];
          }
      }
    }
    __syncthreads();
    // P = 2 (Y2 - 2 X D) on the tile's upper part, consecutive threads along a row
    for (int e = tid; e < PR * TB; e += 256) {
      const int rp = e / TB, cc = e % TB, rr = h * PR + rp;
      if (rr < rmax && cc < cmax && (ti != tj || rr <= cc)) {
        double* t = &T[rp * LDT + cc];
        const double pv = 2.0 * (Ym[(size_t)(i0 + rr) * ld + j0 + cc] - 2.0 * *t);
        Pm[(size_t)(i0 + rr) * n + j0 + cc] = pv;
        *t = pv;
      }
    }
    __syncthreads();
    // mirror: P[j0 + cc][i0 + rr] = tile[rr][cc], consecutive threads along rr
    for (int e = tid; e < PR * TB; e += 256) {
      const int cc = e / PR, rr = h * PR + e % PR;
      if (rr < rmax && cc < cmax && (ti != tj || rr < cc))
        Pm[(size_t)(j0 + cc) * n + i0 + rr] = T[(rr - h * PR) * LDT + cc];
    }
    __syncthreads();
  }
}

// FP64 packed block -> FP32 zero-padded block of the SP2 X layout (rows and columns n .. ld-1 zero).
__global__ void k_pad32(const int* __restrict__ idx, const int* __restrict__ norb,
                        const long long* __restrict__ poff, const long long* __restrict__ xoff,
                        const int* __restrict__ xld, const double* __restrict__ A, float* __restrict__ Ax) {
  const int m = idx[blockIdx.y];
  const int n = norb[m], ld = xld[m];
  const double* src = A + poff[m];
  float* dst = Ax + xoff[m];
  for (int r = blockIdx.x; r < ld; r += gridDim.x)
    for (int c = threadIdx.x; c < ld; c += blockDim.x)
      dst[(size_t)r * ld + c] = (r < n && c < n) ? (float)src[(size_t)r * n + c] : 0.0f;
}

// C = A B for FP64 packed square blocks with B symmetric (B is read by rows: B[k][j] = B[j][k]),
// on bf16x3 MFMA with FP32 accumulation (relative error ~1e-6, enough for a DIIS error vector):
// the operands are rounded to FP32 and split into bf16 hi + lo while they are staged through LDS.
// All 128 x 128 output tiles of each active molecule, same pipeline as k_sp2_bf16x3.
__global__ __launch_bounds__(256) void k_gemm_bf16x3_f64(const int* __restrict__ idx, int ntmax,
                                                         const int* __restrict__ norb,
                                                         const long long* __restrict__ off,
                                                         const double* __restrict__ A, const double* __restrict__ B,
                                                         double* __restrict__ C) {
  __shared__ __attribute__((aligned(16))) __bf16 As[2][2][BT][LDSK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][2][BT][LDSK];
  const int m = idx[blockIdx.y];
  const int n = norb[m];
  const int nt = (n + BT - 1) / BT;
  const int ti = blockIdx.x / ntmax, tj = blockIdx.x % ntmax;
  if (ti >= nt || tj >= nt) return;
  const double* Am = A + off[m];
  const double* Bm = B + off[m];
  double* Cm = C + off[m];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int i0 = ti * BT, j0 = tj * BT;
  floatx16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;
  float4 ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = tid + 256 * q;
      const int row = f >> 3, k = k0 + 4 * (f & 7);
      // unconditional loads from clamped addresses (all of them in flight at once), then masked
      float x[4], y[4];
      const int ra_ = min(i0 + row, n - 1), rb_ = min(j0 + row, n - 1);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kc = min(k + u, n - 1);
        const double va = Am[(size_t)ra_ * n + kc];
        const double vb = Bm[(size_t)rb_ * n + kc];
        x[u] = (i0 + row < n && k + u < n) ? (float)va : 0.0f;
        y[u] = (j0 + row < n && k + u < n) ? (float)vb : 0.0f;
      }
      ra[q] = make_float4(x[0], x[1], x[2], x[3]);
      rb[q] = make_float4(y[0], y[1], y[2], y[3]);
    }
  };
  auto split_store = [&](const float4& v, __bf16 (*hi)[LDSK], __bf16 (*lo)[LDSK], int row, int col) {
    union {
      __bf16 h[4];
      uint2 u;
    } ph, pl;
    const float x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ph.h[u] = (__bf16)x[u];
      pl.h[u] = (__bf16)(x[u] - (float)ph.h[u]);
    }
    *reinterpret_cast<uint2*>(&hi[row][col]) = ph.u;
    *reinterpret_cast<uint2*>(&lo[row][col]) = pl.u;
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = tid + 256 * q;
      const int row = f >> 3, kq = f & 7;
      split_store(ra[q], As[buf][0], As[buf][1], row, 4 * kq);
      split_store(rb[q], Bs[buf][0], Bs[buf][1], row, 4 * kq);
    }
  };
  const int kend = (n + BK - 1) / BK * BK;
  gload(0);
  stage(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      bf16x8 ah[2], al[2], bh[2], bl[2];
      const int kofs = kk + 8 * (lane >> 5);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int r = wr * 64 + a * 32 + (lane & 31);
        ah[a] = *reinterpret_cast<const bf16x8*>(&As[cur][0][r][kofs]);
        al[a] = *reinterpret_cast<const bf16x8*>(&As[cur][1][r][kofs]);
        const int c = wc * 64 + a * 32 + (lane & 31);
        bh[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][0][c][kofs]);
        bl[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][1][c][kofs]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        }
    }
    if (more) stage(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int r = i0 + wr * 64 + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int c = j0 + wc * 64 + b * 32 + (lane & 31);
        if (r < n && c < n) Cm[(size_t)r * n + c] = (double)acc[a][b][v];
      }
}

// Sum the tile partials (fixed order), update the idempotency errors, decide the next branch and
// whether the molecule stops.  e = tr X - ||X||_F^2 = sum_i x_i (1 - x_i) over the eigenvalues x_i
// of X; a molecule stops at iteration k when its trace error is below eps (SP2.py:75-79) and
// e_k >= 8 e_{k-2}^2 with e_{k-2} < 3/16, or at the iteration cap.  In exact arithmetic, once
// e_{k-2} < 3/16 every eigenvalue lies within 1/4 of 0 or 1, and a pair of steps with different
// polynomials maps each deviation d to at most 4 d^2, so e_k <= 4 e_{k-2}^2 (up to the
// d (1 - d) vs d factor, within the margin of 8); the trace test only repeats a polynomial when
// the deviations on the side it grows are below the square of the other side's, where e_k stays
// below that bound too.  So e_k >= 8 e_{k-2}^2 means FP32 rounding has taken over: the expansion
// stops at its first iterate on the rounding floor (the idea of stopping when the error no longer
// follows the exact-arithmetic decrease is that of Kruchinina, Rudberg and Rubensson, J. Chem.
// Theory Comput. 12, 5788 (2016)).  Further FP32 iterations do not help: the trace test then
// chooses on rounding noise and the density error grows again.  The reference's reduced-precision
// rule (trace error not falling, SP2.py:74-78) is not used: occupied deviations (below 1) and
// virtual ones (above 0) cancel in the trace, so the trace error can dip mid-run and stop the
// expansion while eigenvalues are still ~1e-2 from 0 or 1.
__global__ __launch_bounds__(64) void k_sp2_decide(const int* __restrict__ idx, int nact,
                                                   const int* __restrict__ norb, const int* __restrict__ nocc,
                                                   double* __restrict__ st, const double* __restrict__ part,
                                                   int part_stride, int tb, double eps, int max_iter,
                                                   int* __restrict__ ndone) {
  // one wave per molecule: lanes take tile partials t = lane, lane + 64, ... and a fixed shuffle
  // tree sums them (deterministic)
  const int m = idx[blockIdx.x];
  const int lane = threadIdx.x;
  double* s = st + (size_t)m * S_NST;
  if (s[S_DONE] != 0.0) return;  // uniform over the wave
  const int n = norb[m];
  const int nt = (n + tb - 1) / tb;
  const int ntri = nt * (nt + 1) / 2;
  double tr = 0.0, fro = 0.0;
  const double* p = part + (size_t)m * part_stride;
  for (int t = lane; t < ntri; t += 64) {
    tr += p[2 * t];
    fro += p[2 * t + 1];
  }
  tr = wsum(tr);
  fro = wsum(fro);
  if (lane != 0) return;
  const double N = nocc[m];
  s[S_TR] = tr;
  s[S_FRO] = fro;
  s[S_E2] = s[S_E1];
  s[S_E1] = s[S_E0];
  s[S_E0] = fabs(tr - fro);
  s[S_IT] += 1.0;
  s[S_BRANCH] = fabs(fro - N) < fabs(2.0 * tr - fro - N) ? 1.0 : 0.0;
  const bool settled = fabs(tr - N) < eps && s[S_E2] < 0.1875 && s[S_E0] >= 8.0 * s[S_E2] * s[S_E2];
  const bool stop = (settled && s[S_IT] >= 3.0) || s[S_IT] >= max_iter;
  if (stop) {
    s[S_DONE] = 1.0;
    atomicAdd(ndone, 1);
  }
}

// ---------------------------------------------------------------- FP64 GEMM
// C = A * B for symmetric B (B is read by rows, B[k][j] = B[j][k]); A read by rows.  Per active
// molecule, square n x n, row-major with leading dimension n (packed layout) or ld (SP2 layout).
// Tile DT x DT per 256-thread block (4 waves as 2x2), v_mfma_f64_16x16x4_f64, K slab 16 staged
// through registers (prefetched during the MFMAs of the previous slab) and LDS.
// mode bit 0: McWeeny epilogue C <- 2 (3 A - 2 A B) (A = X^2, B = X); bit 1: the product is known
// to be symmetric (X^2, X^2 X): only tiles ti <= tj are computed and mirrored.
// AX / BX: the A / B operand is the final FP32 SP2 matrix X of each molecule (buffer Xa or Xb by
// the parity of its iteration count), widened to FP64 as it is loaded (exact); CPACK: C is written
// to the packed layout (offset coff[m], leading dimension n) instead of the operands' layout.
// DOUT (with AX, symmetric mode): the McWeeny first half; C = X^2 is written on the upper-triangle
// tiles only, and D = C - X (FP32, both triangles) into the SP2 buffer that does not hold X.
template <typename T>
__device__ inline double ld64(const T* p) {
  return (double)*p;
}
template <int DT, bool AX, bool BX, bool CPACK, bool DOUT = false>
__global__ __launch_bounds__(256, 2) void k_dgemm_sym(const int* __restrict__ idx, const int* __restrict__ norb,
                                                   const long long* __restrict__ off, const int* __restrict__ lds,
                                                   const double* __restrict__ A, const double* __restrict__ B,
                                                   double* __restrict__ C, int mode, int ntile_side,
                                                   float* Xa, float* Xb, const double* __restrict__ st,
                                                   const long long* __restrict__ coff) {
  constexpr int WT = DT / 2;    // per-wave tile edge
  constexpr int NB = WT / 16;   // 16x16 MFMA blocks per wave side
  constexpr int NQ = DT * 16 / 256;
  __shared__ double As[DT][17], Bs[DT][17];
  const int m = idx[blockIdx.y];
  const int n = norb[m];
  int ti, tj;
  const bool symout = (mode & 2) != 0;
  if (symout) {
    tri_decode(blockIdx.x, ti, tj);
  } else {
    ti = blockIdx.x / ntile_side;
    tj = blockIdx.x % ntile_side;
  }
  if (ti * DT >= n || tj * DT >= n) return;
  const int ld = lds ? lds[m] : n;
  const size_t o = off[m];
  static_assert(!DOUT || AX, "D = X^2 - X needs X as the A operand");
  const float* Xm = nullptr;
  float* Dm = nullptr;
  if (AX || BX) {
    const bool odd = (((long long)st[(size_t)m * S_NST + S_IT]) & 1) != 0;
    Xm = (odd ? Xb : Xa) + o;
    if (DOUT) Dm = (odd ? Xa : Xb) + o;
  }
  using TA = typename std::conditional<AX, float, double>::type;
  using TB = typename std::conditional<BX, float, double>::type;
  const TA* Am;
  const TB* Bm;
  if constexpr (AX) Am = Xm; else Am = A + o;
  if constexpr (BX) Bm = Xm; else Bm = B + o;
  double* Cm = CPACK ? C + coff[m] : C + o;
  const int ldc = CPACK ? n : ld;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int i0 = ti * DT, j0 = tj * DT;
  doublex4 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = (doublex4){0.0, 0.0, 0.0, 0.0};
  double ra[NQ], rb[NQ];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + 256 * q;
      const int row = e >> 4, gk = k0 + (e & 15);
      const int kc = min(gk, n - 1);
      const double va = ld64(Am + (size_t)min(i0 + row, n - 1) * ld + kc);
      const double vb = ld64(Bm + (size_t)min(j0 + row, n - 1) * ld + kc);
      ra[q] = (i0 + row < n && gk < n) ? va : 0.0;
      rb[q] = (j0 + row < n && gk < n) ? vb : 0.0;
    }
  };
  gload(0);
  for (int k0 = 0; k0 < n; k0 += 16) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + 256 * q;
      As[e >> 4][e & 15] = ra[q];
      Bs[e >> 4][e & 15] = rb[q];
    }
    __syncthreads();
    if (k0 + 16 < n) gload(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      double av[NB], bv[NB];
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        av[a] = As[wr * WT + a * 16 + (lane & 15)][kk + (lane >> 4)];
        bv[a] = Bs[wc * WT + a * 16 + (lane & 15)][kk + (lane >> 4)];
      }
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      // A of the McWeeny form for this 16 x 16 block, loaded before the first use
      double av[4];
      if ((mode & 1) || DOUT) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = min(i0 + wr * WT + a * 16 + 4 * v + (lane >> 4), n - 1);
          const int c = min(j0 + wc * WT + b * 16 + (lane & 15), n - 1);
          av[v] = ld64(Am + (size_t)r * ld + c);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = i0 + wr * WT + a * 16 + 4 * v + (lane >> 4);
        const int c = j0 + wc * WT + b * 16 + (lane & 15);
        if (r < n && c < n) {
          const double x = acc[a][b][v];
          const double y = (mode & 1) ? 2.0 * (3.0 * av[v] - 2.0 * x) : x;
          Cm[(size_t)r * ldc + c] = y;
          if constexpr (DOUT) {
            const float d = (float)(y - av[v]);
            Dm[(size_t)r * ld + c] = d;
            if (ti != tj) Dm[(size_t)c * ld + r] = d;
          } else if (symout && ti != tj) {
            Cm[(size_t)c * ldc + r] = y;
          }
        }
      }
    }
}

inline int st_ok() { return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH; }
inline int eblocks(long long n2) {
  long long b = (n2 + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

}  // namespace

extern "C" int pyseqm_sp2_init(const pyseqm_layout_t* L, const double* F, const int* idx, int nact,
                               const long long* xoff, const int* xld, float* X, double* state, double* red, int nb,
                               void* stream) {
  if (!L || nact < 0 || nb < 1 || nb > PYSEQM_MAX_RED_BLOCKS) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_gersh<<<dim3(nb, nact), 256, 0, s>>>(*L, F, idx, nb, red);
  k_sp2_x0<<<dim3(nb, nact), 256, 0, s>>>(*L, F, idx, nb, xoff, xld, X, red);
  k_sp2_state0<<<(nact + 63) / 64, 64, 0, s>>>(idx, nact, nb, L->mol_nocc, red, state);
  return st_ok();
}

extern "C" int pyseqm_pad_for_eigh(const pyseqm_layout_t* L, const double* F, const int* idx, int nact, int nmax,
                                   double* out, double* red, int nb, void* stream) {
  if (!L || nact < 0 || nmax <= 0 || nb < 1 || nb > PYSEQM_MAX_RED_BLOCKS) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_gersh<<<dim3(nb, nact), 256, 0, s>>>(*L, F, idx, nb, red);
  k_pad_fill<<<dim3(nb, nact), 256, 0, s>>>(*L, F, idx, nb, nmax, red, out);
  return st_ok();
}

extern "C" int pyseqm_sp2_iterate(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                  const long long* xoff, const int* xld, const float* Xin, float* Xout,
                                  double* state, double* part, int part_stride, double eps, int max_iter,
                                  int* ndone, void* stream) {
  if (!L || nact < 0 || nmax <= 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  // 64 x 64 output tiles while the largest active molecule is small (less zero padding in the
  // MFMA work of the many small blocks), 128 x 128 tiles above that
  const int tb = nmax <= kSp2SmallTileMaxN ? 64 : 128;
  const int nt = (nmax + tb - 1) / tb;
  const int ntri = nt * (nt + 1) / 2;
  if (part_stride < nt * (nt + 1)) return PYSEQM_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int grid = 8 * ((nact * ntri + 7) / 8);
  if (tb == 64)
    k_sp2_bf16x3<64><<<grid, 256, 0, s>>>(idx, nact, ntri, L->mol_norb, xoff, xld, Xin, Xout, state, part,
                                          part_stride);
  else
    k_sp2_bf16x3<128><<<grid, 256, 0, s>>>(idx, nact, ntri, L->mol_norb, xoff, xld, Xin, Xout, state, part,
                                           part_stride);
  k_sp2_decide<<<nact, 64, 0, s>>>(idx, nact, L->mol_norb, L->mol_nocc, state, part, part_stride, tb, eps,
                                   max_iter, ndone);
  return st_ok();
}

// `count` SP2 iterations it0 .. it0+count-1 in one call: iteration it reads buffer (it even ? Xa : Xb)
// and writes the other one (the ping-pong of pyseqm_sp2_iterate), with no host work in between.
extern "C" int pyseqm_sp2_iterate_many(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                       const long long* xoff, const int* xld, float* Xa, float* Xb, double* state,
                                       double* part, int part_stride, double eps, int max_iter, int* ndone, int it0,
                                       int count, void* stream) {
  if (count < 0 || it0 < 0) return PYSEQM_EINVAL;
  for (int k = 0; k < count; ++k) {
    const bool even = ((it0 + k) & 1) == 0;
    const int rc = pyseqm_sp2_iterate(L, idx, nact, nmax, xoff, xld, even ? Xa : Xb, even ? Xb : Xa, state, part,
                                      part_stride, eps, max_iter, ndone, stream);
    if (rc != PYSEQM_OK) return rc;
  }
  return PYSEQM_OK;
}

extern "C" int pyseqm_diis_commutator_bf16x3(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                             const double* F, const double* P, const long long* xoff, const int* xld,
                                             float* Fx, float* Px, double* Cout, double* diis_err, double* tmax,
                                             int tstride, void* stream) {
  if (!L || nact < 0 || nmax <= 0) return PYSEQM_EINVAL;
  const int tb = nmax <= kSp2SmallTileMaxN ? 64 : 128;
  const int nt = (nmax + tb - 1) / tb;
  const int ntri = nt * (nt + 1) / 2;
  if (tstride < ntri) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  const int rows = (nmax + 31) / 32 * 32;
  const dim3 pgrid(rows < 2048 ? rows : 2048, nact);
  k_pad32<<<pgrid, 256, 0, s>>>(idx, L->mol_norb, L->mol_mat_off, xoff, xld, F, Fx);
  k_pad32<<<pgrid, 256, 0, s>>>(idx, L->mol_norb, L->mol_mat_off, xoff, xld, P, Px);
  const int grid = 8 * ((nact * ntri + 7) / 8);
  if (tb == 64)
    k_comm_bf16x3<64><<<grid, 256, 0, s>>>(idx, nact, ntri, L->mol_norb, xoff, xld, Fx, Px, L->mol_mat_off, Cout,
                                           tmax, tstride);
  else
    k_comm_bf16x3<128><<<grid, 256, 0, s>>>(idx, nact, ntri, L->mol_norb, xoff, xld, Fx, Px, L->mol_mat_off, Cout,
                                            tmax, tstride);
  k_comm_final<<<nact, 64, 0, s>>>(idx, L->mol_norb, tmax, tstride, tb, diis_err);
  return st_ok();
}

extern "C" int pyseqm_gemm_bf16x3(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* A,
                                  const double* B, double* C, void* stream) {
  if (!L || nact < 0 || nmax <= 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const int nt = (nmax + BT - 1) / BT;
  k_gemm_bf16x3_f64<<<dim3(nt * nt, nact), 256, 0, (hipStream_t)stream>>>(idx, nt, L->mol_norb, L->mol_mat_off, A,
                                                                           B, C);
  return st_ok();
}

extern "C" int pyseqm_dgemm_sym(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                const long long* off, const int* lds, const double* A, const double* B, double* C,
                                int mode, void* stream) {
  if (!L || nact < 0 || nmax <= 0 || mode < 0 || mode > 3) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  const bool symout = (mode & 2) != 0;
  if (nmax <= 256) {
    const int nt = (nmax + 63) / 64;
    k_dgemm_sym<64, false, false, false><<<dim3(symout ? nt * (nt + 1) / 2 : nt * nt, nact), 256, 0, s>>>(
        idx, L->mol_norb, off, lds, A, B, C, mode, nt, nullptr, nullptr, nullptr, nullptr);
  } else {
    const int nt = (nmax + 127) / 128;
    k_dgemm_sym<128, false, false, false><<<dim3(symout ? nt * (nt + 1) / 2 : nt * nt, nact), 256, 0, s>>>(
        idx, L->mol_norb, off, lds, A, B, C, mode, nt, nullptr, nullptr, nullptr, nullptr);
  }
  return st_ok();
}

// McWeeny purification of the final SP2 matrix: Y2 = X X on the FP64 MFMA GEMM (upper-triangle tiles,
// X widened exactly as it is loaded) with D = Y2 - X into the other SP2 buffer, then
// P = 2 (Y2 - 2 X D) (= 2 (3 X^2 - 2 X^3)) on bf16x3 MFMA into the packed layout.
template <int DT>
static void launch_mcweeny(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* xoff,
                           const int* xld, const double* state, float* Xa, float* Xb, double* Y2, double* P,
                           hipStream_t s) {
  const int nt = (nmax + DT - 1) / DT;
  const int ntri = nt * (nt + 1) / 2;
  k_dgemm_sym<DT, true, true, false, true><<<dim3(ntri, nact), 256, 0, s>>>(
      idx, L->mol_norb, xoff, xld, nullptr, nullptr, Y2, 2, nt, Xa, Xb, state, nullptr);
  k_mcw_bf16x3<DT><<<8 * ((nact * ntri + 7) / 8), 256, 0, s>>>(idx, nact, ntri, L->mol_norb, xoff, xld, Xa, Xb,
                                                                state, Y2, L->mol_mat_off, P);
}

extern "C" int pyseqm_sp2_mcweeny(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                  const long long* xoff, const int* xld, const double* state, float* Xa,
                                  float* Xb, double* Y2, double* P, void* stream) {
  if (!L || nact < 0 || nmax <= 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  if (nmax <= 256)
    launch_mcweeny<64>(L, idx, nact, nmax, xoff, xld, state, Xa, Xb, Y2, P, s);
  else
    launch_mcweeny<128>(L, idx, nact, nmax, xoff, xld, state, Xa, Xb, Y2, P, s);
  return st_ok();
}

