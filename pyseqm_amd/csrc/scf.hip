// SCF-loop kernels on the packed (ragged) batch, FP64, gfx950.
//
// All kernels work on a list of active molecules idx[0..nact) (device array) so that converged
// molecules cost nothing; grid.y indexes that list.  They restate, per molecule:
//   density-step plumbing of sym_eig_trunc (diag.py:110-241): Gershgorin-shifted padding of the
//     packed Fock matrix for a batched eigensolver and P = 2 C_occ C_occ^T,
//   mixing (scf_loop.py:266-267, 1052-1056) and the 4-way convergence test get_error
//     (scf_loop.py:106-147),
//   the Pulay/DIIS residual FP - PF, its max-norm, the B-matrix row and the extrapolated Fock
//     matrix (scf_loop.py:997-1035).
#include "common.h"

using namespace pyseqm;

namespace {

constexpr int kEl = 256;  // threads of the element-wise / reduction kernels

__device__ inline double block_sum(double v, double* sh) {
  v = wave64(v);  // DPP / permlane swaps, the pairings of the xor tree (every thread active here)
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  __syncthreads();
  return s;  // valid in thread 0
}

__device__ inline double wave_max(double v) {
  for (int o = 1; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__device__ inline double block_max(double v, double* sh) {
  v = wave_max(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = -1.0e308;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s = fmax(s, sh[k]);
  __syncthreads();
  return s;
}

// P = 2 sum_{k < nocc} V[:,k] V[:,k]^T (diag.py:230-232) from the padded eigenvectors (columns),
// 16x16 output tile per block, LDS-staged rows of V.
__global__ void k_density_from_evecs(pyseqm_layout_t L, const double* __restrict__ V, int nmax,
                                     const int* __restrict__ idx, double* __restrict__ P) {
  __shared__ double Va[16][33], Vb[16][33];
  const int k = blockIdx.y;
  const int m = idx[k];
  // V holds nmax x nmax per listed block: reads stay inside it even if a caller breaks the
  // nmax >= norb precondition (the host wrappers check it)
  const int n = min(L.mol_norb[m], nmax);
  const int nt = (nmax + 15) / 16;
  const int ti = blockIdx.x / nt, tj = blockIdx.x % nt;
  if (ti * 16 >= n || tj * 16 >= n) return;
  const int nocc = min(L.mol_nocc[m], nmax);
  const double* Vm = V + (size_t)k * nmax * nmax;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int r = ti * 16 + ty, c = tj * 16 + tx;
  double s = 0.0;
  for (int k0 = 0; k0 < nocc; k0 += 32) {
    for (int e = threadIdx.x; e < 16 * 32; e += 256) {
      const int rr = e / 32, kk = e % 32;
      const int gi = ti * 16 + rr, gj = tj * 16 + rr, gk = k0 + kk;
      Va[rr][kk] = (gi < n && gk < nocc) ? Vm[(size_t)gi * nmax + gk] : 0.0;
      Vb[rr][kk] = (gj < n && gk < nocc) ? Vm[(size_t)gj * nmax + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; ++kk) s += Va[ty][kk] * Vb[tx][kk];
    __syncthreads();
  }
  if (r < n && c < n) P[L.mol_mat_off[m] + (size_t)r * n + c] = 2.0 * s;
}

// Pold = P; P = a*P + (1-a)*Pnew on the active molecules.
// Element-wise kernels take kElPerThread elements per thread, loads first (a load-then-store loop
// would wait out one memory round trip per element), so a molecule needs n^2 / (kEl kElPerThread)
// blocks and small molecules do not dispatch a large molecule's block count of empty blocks.
constexpr int kElPerThread = 4;

__global__ void k_mix(pyseqm_layout_t L, const int* __restrict__ idx, double a, double* __restrict__ P,
                      const double* __restrict__ Pnew, double* __restrict__ Pold) {
  const int m = idx[blockIdx.y];
  const size_t n2 = (size_t)L.mol_norb[m] * L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t e0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < n2; e0 += kElPerThread * step) {
    double p[kElPerThread], q[kElPerThread];
#pragma unroll
    for (int u = 0; u < kElPerThread; ++u) {
      const size_t e = e0 + u * step;
      p[u] = e < n2 ? P[off + e] : 0.0;
      q[u] = e < n2 ? Pnew[off + e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kElPerThread; ++u) {
      const size_t e = e0 + u * step;
      if (e < n2) {
        if (Pold) Pold[off + e] = p[u];
        P[off + e] = a * p[u] + (1.0 - a) * q[u];
      }
    }
  }
}

// ------------------------------------------------------------------ adaptive mixing
// scf_converger [1] (adaptive_mix / compute_fac, scf_loop.py:350-420) on the active molecules:
//   every third iteration P = (1 + FAC) Pnew - FAC P off the diagonal, FAC = sqrt(|dD|^2 / |d2D|^2)
//   from the diagonals (D = diag P, d2D = Dnew - 2 D + D_old2; FAC = 0 unless d2D != 0 and
//   |dD|^2 < 100 |d2D|^2), else P = Pnew; the diagonal steps by at most DAMP = 0.05 after
//   iteration 4 (else by Dnew - D + FAC (Dnew - D)), clamped to [0, 2], then renormalised to the
//   trace of Dnew: up to 20 passes that scale every active diagonal by SUM0 / SUM2, pin values
//   above 2 at 2 and move them out of SUM0.  The reference stops the passes when every active
//   molecule is renormalised (|SUM0 / SUM2 - 1| <= 1e-5, or SUM2 <= 1e-3), so a molecule keeps
//   being scaled while any other one is not done.  Each molecule's sequence of passes depends on
//   its own diagonal only: k_amix_passes runs the 20 passes of every molecule and ANDs into `mask`
//   the passes at which it is done; the batch stops at the lowest pass all molecules agree on, and
//   k_amix_diag applies exactly that many passes (the same arithmetic, so the same values).
constexpr int kMixPasses = 20;
constexpr unsigned kMixAllPasses = (1u << kMixPasses) - 1u;

// Diagonal after damping / clamping (left in di[0..n)) and FAC, SUM0 of one molecule.
__device__ inline void amix_start(const double* __restrict__ Pprev, const double* __restrict__ Pcur,
                                  const double* __restrict__ old2, int n, int k, double* di, double& fac,
                                  double& sum0, double* sh) {
  const bool third = k % 3 == 0;
  const double damp = k > 4 ? 0.05 : 1.0e10;
  fac = 0.0;
  if (third) {
    double num = 0.0, den = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const double dp = Pprev[(size_t)r * (n + 1)], dc = Pcur[(size_t)r * (n + 1)];
      const double d1 = dc - dp, d2 = dc - 2.0 * dp + old2[r];
      num += d1 * d1;
      den += d2 * d2;
    }
    num = block_sum(num, sh);
    den = block_sum(den, sh);
    if (threadIdx.x == 0) sh[8] = (den > 0.0 && num < 100.0 * den) ? sqrt(num / den) : 0.0;
    __syncthreads();
    fac = sh[8];
    __syncthreads();
  }
  double s0 = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double dp = Pprev[(size_t)r * (n + 1)], dc = Pcur[(size_t)r * (n + 1)];
    const double delta = dc - dp;
    double v = fabs(delta) > damp ? dp + (delta > 0.0 ? damp : -damp) : dc + fac * delta;
    di[r] = fmin(fmax(v, 0.0), 2.0);
    s0 += dc;
  }
  sum0 = block_sum(s0, sh);  // broadcast to every thread (block_sum ends with a barrier)
  if (threadIdx.x == 0) sh[9] = sum0;
  __syncthreads();
  sum0 = sh[9];
  __syncthreads();
}

// One renormalisation pass; returns whether the molecule was already done at its start.
__device__ inline bool amix_pass(double* di, int n, double& sum0, double* sh) {
  double s2 = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) s2 += di[r];
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) sh[10] = s2;
  __syncthreads();
  s2 = sh[10];
  const bool large = s2 > 1.0e-3;
  const double sum3 = large ? sum0 / s2 : 0.0;
  const bool done = !large || fabs(sum3 - 1.0) <= 1.0e-5;
  double nfull = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double sc = fmax(di[r] * sum3, 0.0);
    const bool full = sc > 2.0;
    di[r] = full ? 2.0 : sc;
    nfull += full ? 1.0 : 0.0;
  }
  nfull = block_sum(nfull, sh);
  if (threadIdx.x == 0) sh[11] = nfull;
  __syncthreads();
  sum0 -= 2.0 * sh[11];
  __syncthreads();
  return done;
}

__global__ void k_amix_init(unsigned* mask) {
  if (threadIdx.x == 0) *mask = kMixAllPasses;
}

__global__ __launch_bounds__(kEl) void k_amix_passes(pyseqm_layout_t L, const int* __restrict__ idx, int k,
                                                     const double* __restrict__ P, const double* __restrict__ Pnew,
                                                     const double* __restrict__ old2,
                                                     const long long* __restrict__ orb_off,
                                                     double* __restrict__ fac_out, unsigned* __restrict__ mask) {
  extern __shared__ double di[];
  __shared__ double sh[16];
  const int m = idx[blockIdx.x];
  const int n = L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  double fac, sum0;
  amix_start(P + off, Pnew + off, old2 + orb_off[m], n, k, di, fac, sum0, sh);
  unsigned bits = 0;
  for (int q = 0; q < kMixPasses; ++q)
    if (amix_pass(di, n, sum0, sh)) bits |= 1u << q;
  if (threadIdx.x == 0) {
    fac_out[m] = fac;
    atomicAnd(mask, bits);
  }
}

// Off-diagonal mixing (and Pold = P) over every element; k_amix_diag then rewrites the diagonal.
__global__ void k_amix_elems(pyseqm_layout_t L, const int* __restrict__ idx, int third,
                             const double* __restrict__ fac, double* __restrict__ P, const double* __restrict__ Pnew,
                             double* __restrict__ Pold) {
  const int m = idx[blockIdx.y];
  const size_t n2 = (size_t)L.mol_norb[m] * L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  const double f = third ? fac[m] : 0.0;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (size_t e0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < n2; e0 += kElPerThread * step) {
    double p[kElPerThread], q[kElPerThread];
#pragma unroll
    for (int u = 0; u < kElPerThread; ++u) {
      const size_t e = e0 + u * step;
      p[u] = e < n2 ? P[off + e] : 0.0;
      q[u] = e < n2 ? Pnew[off + e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kElPerThread; ++u) {
      const size_t e = e0 + u * step;
      if (e < n2) {
        Pold[off + e] = p[u];
        P[off + e] = third ? (1.0 + f) * q[u] - f * p[u] : q[u];
      }
    }
  }
}

__global__ __launch_bounds__(kEl) void k_amix_diag(pyseqm_layout_t L, const int* __restrict__ idx, int k,
                                                   double* __restrict__ P, const double* __restrict__ Pnew,
                                                   const double* __restrict__ Pold, double* __restrict__ old2,
                                                   const long long* __restrict__ orb_off,
                                                   const unsigned* __restrict__ mask) {
  extern __shared__ double di[];
  __shared__ double sh[16];
  const int m = idx[blockIdx.x];
  const int n = L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  double* o2 = old2 + orb_off[m];
  double fac, sum0;
  amix_start(Pold + off, Pnew + off, o2, n, k, di, fac, sum0, sh);  // Pold holds the unmixed P
  const unsigned agree = *mask & kMixAllPasses;
  const int npass = agree ? __builtin_ctz(agree) : kMixPasses;
  for (int q = 0; q < npass; ++q) amix_pass(di, n, sum0, sh);
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    P[off + (size_t)r * (n + 1)] = di[r];
    o2[r] = Pold[off + (size_t)r * (n + 1)];
  }
}

// Convergence test of get_error (scf_loop.py:106-147): |dE| <= eps, max|FP-PF| <= diis_tol
// (50 eps in the reference; the mixed-precision density step passes its precision floor),
// ||dP||_F / (4 natoms) <= 2 eps, max|dP| <= 15 eps.  Two stages: per-block partial sums of
// dP^2 and max|dP| over nb blocks per molecule, then one thread per molecule (fixed order).
__global__ void k_check_partial(pyseqm_layout_t L, const int* __restrict__ idx, int nb,
                                const double* __restrict__ P, const double* __restrict__ Pold,
                                double* __restrict__ red) {
  __shared__ double sh[kEl / 64];
  const int k = blockIdx.y;
  const int m = idx[k];
  const size_t n2 = (size_t)L.mol_norb[m] * L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  double ss = 0.0, mx = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)nb * blockDim.x) {
    const double d = P[off + e] - Pold[off + e];
    ss += d * d;
    mx = fmax(mx, fabs(d));
  }
  const double s2 = block_sum(ss, sh);
  const double m2 = block_max(mx, sh);
  if (threadIdx.x == 0) {
    red[((size_t)k * nb + blockIdx.x) * 2] = s2;
    red[((size_t)k * nb + blockIdx.x) * 2 + 1] = m2;
  }
}

__global__ __launch_bounds__(64) void k_check_final(pyseqm_layout_t L, const int* __restrict__ idx, int nact, int nb,
                                                   const double* __restrict__ red, const double* __restrict__ e_new,
                                                   double* __restrict__ e_old, const double* __restrict__ diis_err,
                                                   double eps, double diis_tol, unsigned char* __restrict__ notconv,
                                                   double* __restrict__ metrics) {
  // one wave per molecule: lanes take block partials b = lane, lane + 64, ... (fixed order)
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  const int m = idx[k];
  double s2 = 0.0, m2 = 0.0;
  for (int b = lane; b < nb; b += 64) {
    s2 += red[((size_t)k * nb + b) * 2];
    m2 = fmax(m2, red[((size_t)k * nb + b) * 2 + 1]);
  }
  s2 = wave_sum(s2);
  m2 = wave_max(m2);
  if (lane != 0) return;
  const int nat = L.mol_atom_off[m + 1] - L.mol_atom_off[m];
  const double dm = sqrt(s2) / (4.0 * nat);
  const double de = e_new[m] - e_old[m];
  const double dd = diis_err ? diis_err[m] : 0.0;
  const bool bad = fabs(de) > eps || dd > diis_tol || dm > 2.0 * eps || m2 > 15.0 * eps;
  notconv[m] = bad ? 1 : 0;
  if (bad) e_old[m] = e_new[m];
  if (metrics) {
    metrics[4 * m + 0] = de;
    metrics[4 * m + 1] = dd;
    metrics[4 * m + 2] = dm;
    metrics[4 * m + 3] = m2;
  }
}

// Ragged FP64 GEMM C = A B per active molecule (n x n), 64x64 output tile per 256-thread block,
// 4x4 outputs per thread, K staged through LDS in slabs of 16.
__global__ __launch_bounds__(256) void k_gemm64(pyseqm_layout_t L, const int* __restrict__ idx, int nmax,
                                                const double* __restrict__ A, const double* __restrict__ B,
                                                double* __restrict__ C) {
  __shared__ double As[16][64 + 1], Bs[16][64 + 1];
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const int nt = (nmax + 63) / 64;
  const int ti = blockIdx.x / nt, tj = blockIdx.x % nt;
  if (ti * 64 >= n || tj * 64 >= n) return;
  const size_t off = L.mol_mat_off[m];
  const double* Am = A + off;
  const double* Bm = B + off;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < n; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e & 15, rr = e >> 4;  // A tile: rows rr, cols k0+kk
      const int gi = ti * 64 + rr, gk = k0 + kk;
      As[kk][rr] = (gi < n && gk < n) ? Am[(size_t)gi * n + gk] : 0.0;
      const int cc = e & 63, kb = e >> 6;  // B tile: rows k0+kb, cols cc
      const int gkb = k0 + kb, gj = tj * 64 + cc;
      Bs[kb][cc] = (gkb < n && gj < n) ? Bm[(size_t)gkb * n + gj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[kk][ty + 16 * u];
        b[u] = Bs[kk][tx + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int gi = ti * 64 + ty + 16 * u, gj = tj * 64 + tx + 16 * v;
      if (gi < n && gj < n) C[off + (size_t)gi * n + gj] = acc[u][v];
    }
}

// C = X - X^T (X = F P, so C = FP - PF) into the ring slot, with per-block partial max|C|.
// C = X - X^T and per-block max|C| over the active blocks.  Work unit: a pair of mirrored 32 x 32
// tiles (I, J), I <= J, of one molecule; both are read as coalesced rows into LDS and both output
// tiles are written as rows (C_JI = -C_IJ^T), so no element of X is read down a column of HBM.
__global__ __launch_bounds__(kEl) void k_antisym(pyseqm_layout_t L, const int* __restrict__ idx, int nb,
                                                 const double* __restrict__ X, double* __restrict__ Cout,
                                                 double* __restrict__ red) {
  __shared__ double A[32][33], B[32][33];
  __shared__ double sh[kEl / 64];
  const int k = blockIdx.y;
  const int m = idx[k];
  const int n = L.mol_norb[m];
  const double* Xm = X + L.mol_mat_off[m];
  double* Cm = Cout + L.mol_mat_off[m];
  const int T = (n + 31) / 32;
  const int ntri = T * (T + 1) / 2;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows of 32 per pass
  double mx = 0.0;
  for (int t = blockIdx.x; t < ntri; t += nb) {
    int tj = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((tj + 1) * (tj + 2) / 2 <= t) ++tj;
    while (tj * (tj + 1) / 2 > t) --tj;
    const int ti = t - tj * (tj + 1) / 2;
    const int i0 = ti * 32, j0 = tj * 32;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = ty + 8 * q;
      // unconditional loads from clamped addresses, masked afterwards
      const double va = Xm[(size_t)min(i0 + rr, n - 1) * n + min(j0 + tx, n - 1)];
      const double vb = Xm[(size_t)min(j0 + rr, n - 1) * n + min(i0 + tx, n - 1)];
      A[rr][tx] = (i0 + rr < n && j0 + tx < n) ? va : 0.0;
      B[rr][tx] = (j0 + rr < n && i0 + tx < n) ? vb : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = ty + 8 * q;
      if (i0 + rr < n && j0 + tx < n) {
        const double v = A[rr][tx] - B[tx][rr];  // X[r][c] - X[c][r]
        Cm[(size_t)(i0 + rr) * n + j0 + tx] = v;
        mx = fmax(mx, fabs(v));
      }
      if (ti != tj && j0 + rr < n && i0 + tx < n) {
        const double v = B[rr][tx] - A[tx][rr];
        Cm[(size_t)(j0 + rr) * n + i0 + tx] = v;
        mx = fmax(mx, fabs(v));
      }
    }
    __syncthreads();
  }
  const double r = block_max(mx, sh);
  if (threadIdx.x == 0) red[(size_t)k * nb + blockIdx.x] = r;
}

__global__ __launch_bounds__(64) void k_antisym_final(const int* __restrict__ idx, int nact, int nb,
                                                     const double* __restrict__ red, double* __restrict__ diis_err) {
  const int k = blockIdx.x;
  double mx = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) mx = fmax(mx, red[(size_t)k * nb + b]);
  mx = wave_max(mx);
  if (threadIdx.x == 0) diis_err[idx[k]] = mx;
}

// B[m][slot][j] = <C_slot, C_j> over the strict upper triangle (= 0.5 * full Frobenius product: C is
// antisymmetric with a zero diagonal), j = 0..count-1 (scf_loop.py:1003-1013).  One wave per row r
// (lanes over the contiguous segment c > r of the row), so only half of each ring entry is read;
// each wave forms all `count` products from one read of C_slot; per-block partials, then a
// fixed-order final sum.
__global__ __launch_bounds__(kEl) void k_diis_dots(pyseqm_layout_t L, const int* __restrict__ idx, int nb,
                                                   const double* __restrict__ ring, size_t ring_stride, int slot,
                                                   int count, double* __restrict__ red) {
  __shared__ double sh[16][kEl / 64];
  const int k = blockIdx.y;
  const int m = idx[k];
  const int n = L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  const double* a = ring + (size_t)slot * ring_stride + off;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  for (int r = blockIdx.x * (kEl / 64) + wave; r < n - 1; r += nb * (kEl / 64)) {
    const size_t row = (size_t)r * n;
    for (int c = r + 1 + lane; c < n; c += 64) {
      const double av = a[row + c];
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j < count) acc[j] += av * ring[(size_t)j * ring_stride + off + row + c];
    }
  }
  // all `count` products reduced together: wave sums on DPP / permlane swaps (chains interleaved),
  // one barrier, then thread j adds the wave totals in wave order
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (j < count) {  // uniform
      const double w = wave64(acc[j]);
      if (lane == 0) sh[j][wave] = w;
    }
  __syncthreads();
  if ((int)threadIdx.x < count) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < kEl / 64; ++w) r += sh[threadIdx.x][w];
    red[((size_t)k * nb + blockIdx.x) * 16 + threadIdx.x] = r;
  }
}

__global__ __launch_bounds__(64) void k_diis_dots_final(const int* __restrict__ idx, int nact, int nb, int slot,
                                                       int count, int nfock, const double* __restrict__ red,
                                                       double* __restrict__ Bmat) {
  // one wave per (molecule, history entry): lanes take block partials b = lane, lane + 64, ...
  const int k = blockIdx.x / count, j = blockIdx.x % count;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) s += red[((size_t)k * nb + b) * 16 + j];
  s = wave_sum(s);
  if (threadIdx.x == 0) Bmat[((size_t)idx[k] * (nfock + 1) + slot) * (nfock + 1) + j] = s;
}

// DIIS coefficients from the bordered B matrix (scf_loop.py:1015-1031): the lower triangle of
// EMAT[m][0..cF][0..cF] (mirrored), its cF x cF block divided by B[slot][slot], eigendecomposed by
// cyclic Jacobi in FP64 (one 64-lane block per molecule, matrix and vectors in LDS, lane i owns
// row / column i), then c_k = -sum_i Q[k][i] q_i Q[cF][i] with q_i = 1/lambda_i for |lambda_i| >
// 1e-13 (else 0), and reset[m] = (max|lambda| / min|lambda| > 1e7).
constexpr int kDiisMax = 16;

__global__ __launch_bounds__(64) void k_diis_solve(const int* __restrict__ idx, int nfock, int cF, int slot,
                                                   const double* __restrict__ EMAT, double* __restrict__ coef,
                                                   unsigned char* __restrict__ reset) {
  __shared__ double A[kDiisMax][kDiisMax + 1], V[kDiisMax][kDiisMax + 1];
  const int m = idx[blockIdx.x];
  const int lane = threadIdx.x;
  const int n = cF + 1, nf1 = nfock + 1;
  const double* E = EMAT + (size_t)m * nf1 * nf1;
  const double den = fmax(E[slot * nf1 + slot], 1.0e-15);
  for (int e = lane; e < n * n; e += 64) {
    const int i = e / n, j = e % n;
    double v = i >= j ? E[i * nf1 + j] : E[j * nf1 + i];
    if (i < cF && j < cF) v /= den;
    A[i][j] = v;
    V[i][j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a = A[i][j];
        tot += a * a;
        if (j > i) off += a * a;
      }
    if (off <= 1.0e-34 * tot) break;  // the same decision in every lane
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p][q];
        if (fabs(apq) < 1.0e-300) continue;
        const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = fabs(tau) < 1.0e150 ? (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau))
                                             : 0.5 / tau;
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
        __syncthreads();
        if (lane < n) {
          const double ap = A[lane][p], aq = A[lane][q];
          A[lane][p] = c * ap - sn * aq;
          A[lane][q] = sn * ap + c * aq;
        }
        __syncthreads();
        if (lane < n) {
          const double ap = A[p][lane], aq = A[q][lane];
          A[p][lane] = c * ap - sn * aq;
          A[q][lane] = sn * ap + c * aq;
          const double vp = V[lane][p], vq = V[lane][q];
          V[lane][p] = c * vp - sn * vq;
          V[lane][q] = sn * vp + c * vq;
        }
        __syncthreads();
      }
  }
  if (lane < cF) {
    double ck = 0.0;
    for (int i = 0; i < n; ++i) {
      const double li = A[i][i];
      const double qi = fabs(li) > 1.0e-13 ? 1.0 / li : 0.0;
      ck -= V[lane][i] * qi * V[cF][i];
    }
    coef[(size_t)m * nfock + lane] = ck;
  }
  if (lane == 0) {
    double lo = 1.0e308, hi = 0.0;
    for (int i = 0; i < n; ++i) {
      lo = fmin(lo, fabs(A[i][i]));
      hi = fmax(hi, fabs(A[i][i]));
    }
    reset[m] = hi > 1.0e7 * lo ? 1 : 0;
  }
}

// F = sum_j c[m][j] F_j over the stored Fock matrices (scf_loop.py:1033-1035).
// F = sum_j coef[j] Fring[j] on the active blocks (scf_loop.py:1033-1035).  The ring entries are
// symmetric and only their upper triangles are kept current (pyseqm_copy_active copies the upper
// triangle): each block takes 32 x 32 tile pairs (I <= J), forms the upper tile from coalesced row
// reads, and writes it and its transpose (through LDS) as rows of F.
__global__ __launch_bounds__(kEl) void k_diis_combine(pyseqm_layout_t L, const int* __restrict__ idx, int nb,
                                                      const double* __restrict__ ring, size_t ring_stride, int count,
                                                      const double* __restrict__ coef, int cstride,
                                                      double* __restrict__ F) {
  __shared__ double S[32][33];
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  const double* c = coef + (size_t)m * cstride;
  const int T = (n + 31) / 32;
  const int ntri = T * (T + 1) / 2;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int t = blockIdx.x; t < ntri; t += nb) {
    int tj = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((tj + 1) * (tj + 2) / 2 <= t) ++tj;
    while (tj * (tj + 1) / 2 > t) --tj;
    const int ti = t - tj * (tj + 1) / 2;
    const int i0 = ti * 32, j0 = tj * 32;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = ty + 8 * q;
      const int r = i0 + rr, cc = j0 + tx;
      // element (r, cc) of the symmetric matrix, read from the upper triangle
      const int ur = min(r, cc), uc = max(r, cc);
      double v = 0.0;
      if (r < n && cc < n) {
        const size_t e = off + (size_t)ur * n + uc;
        for (int j = 0; j < count; ++j) v += c[j] * ring[(size_t)j * ring_stride + e];
        F[off + (size_t)r * n + cc] = v;
      }
      S[rr][tx] = v;
    }
    __syncthreads();
    if (ti != tj) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rr = ty + 8 * q;
        if (j0 + rr < n && i0 + tx < n) F[off + (size_t)(j0 + rr) * n + i0 + tx] = S[tx][rr];
      }
    }
    __syncthreads();
  }
}

// dst = src on the upper triangle (c >= r) of the active molecules' blocks: one wave per row.
__global__ __launch_bounds__(kEl) void k_copy_active(pyseqm_layout_t L, const int* __restrict__ idx, int nb,
                                                     const double* __restrict__ src, double* __restrict__ dst) {
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = blockIdx.x * (kEl / 64) + wave; r < n; r += nb * (kEl / 64)) {
    const size_t row = off + (size_t)r * n;
    for (int c = r + lane; c < n; c += 64) dst[row + c] = src[row + c];
  }
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

inline int elem_blocks(int nmax) {
  const long long n2 = (long long)nmax * nmax;
  int b = ceil_div(n2, (long long)kEl * kElPerThread);
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

inline int status() { return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH; }

// blocks per molecule of the one-wave-per-row kernels (4 rows per wave, 16 per block) and of the
// 32 x 32 tile-pair kernels (2 pairs per block), capped (the kernels loop): sized by the largest
// block, so fewer and fuller blocks keep small molecules from dispatching many empty ones
inline int row_blocks(int nmax) {
  const int b = ceil_div(nmax, 4 * (kEl / 64));
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}
inline int tile_pair_blocks(int nmax) {
  const long long t = ceil_div(nmax, 32);
  const long long b = ceil_div(t * (t + 1) / 2, 2);
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

}  // namespace

extern "C" int pyseqm_density_from_evecs(const pyseqm_layout_t* L, const double* V, int nmax, const int* idx,
                                         int nact, double* P, void* stream) {
  if (!L || nact < 0 || nmax <= 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const int nt = ceil_div(nmax, 16);
  k_density_from_evecs<<<dim3(nt * nt, nact), 256, 0, (hipStream_t)stream>>>(*L, V, nmax, idx, P);
  return status();
}

extern "C" int pyseqm_mix(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, double alpha, double* P,
                          const double* Pnew, double* Pold, void* stream) {
  if (!L || nact < 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  k_mix<<<dim3(elem_blocks(nmax), nact), kEl, 0, (hipStream_t)stream>>>(*L, idx, alpha, P, Pnew, Pold);
  return status();
}

extern "C" int pyseqm_adaptive_mix(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int k, double* P,
                                   const double* Pnew, double* Pold, double* old2, const long long* orb_off,
                                   double* fac, unsigned* mask, void* stream) {
  if (!L || nact < 0 || k < 0 || nmax < 0 || nmax > 8192) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  const size_t lds = (size_t)nmax * sizeof(double);
  k_amix_init<<<1, 64, 0, s>>>(mask);
  k_amix_passes<<<nact, kEl, lds, s>>>(*L, idx, k, P, Pnew, old2, orb_off, fac, mask);
  k_amix_elems<<<dim3(elem_blocks(nmax), nact), kEl, 0, s>>>(*L, idx, k % 3 == 0, fac, P, Pnew, Pold);
  k_amix_diag<<<nact, kEl, lds, s>>>(*L, idx, k, P, Pnew, Pold, old2, orb_off, mask);
  return status();
}

extern "C" int pyseqm_copy_active(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* src,
                                  double* dst, void* stream) {
  if (!L || nact < 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const int nb = row_blocks(nmax);
  k_copy_active<<<dim3(nb, nact), kEl, 0, (hipStream_t)stream>>>(*L, idx, nb, src, dst);
  return status();
}

extern "C" int pyseqm_scf_check(const pyseqm_layout_t* L, const int* idx, int nact, const double* P,
                                const double* Pold, const double* e_new, double* e_old, const double* diis_err,
                                double eps, double diis_tol, unsigned char* notconv, double* metrics,
                                double* red, int nb, void* stream) {
  if (!L || nact < 0 || nb < 1 || nb > PYSEQM_MAX_RED_BLOCKS) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_check_partial<<<dim3(nb, nact), kEl, 0, s>>>(*L, idx, nb, P, Pold, red);
  k_check_final<<<nact, 64, 0, s>>>(*L, idx, nact, nb, red, e_new, e_old, diis_err, eps, diis_tol,
                                                  notconv, metrics);
  return status();
}

extern "C" int pyseqm_gemm(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* A,
                           const double* B, double* C, void* stream) {
  if (!L || nact < 0 || nmax <= 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const int nt = ceil_div(nmax, 64);
  k_gemm64<<<dim3(nt * nt, nact), 256, 0, (hipStream_t)stream>>>(*L, idx, nmax, A, B, C);
  return status();
}

extern "C" int pyseqm_dgemm_sym(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                const long long* off, const int* lds, const double* A, const double* B, double* C,
                                int mode, void* stream);

extern "C" int pyseqm_gemm_bf16x3(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* A,
                                  const double* B, double* C, void* stream);

// X = F P on the FP64 MFMA GEMM (P symmetric), or on the bf16x3 MFMA GEMM when mixed != 0 (the
// mixed-precision density step, whose own error floor is far above FP32 rounding), then
// C = X - X^T and max|C|.
extern "C" int pyseqm_diis_residual(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* F,
                                    const double* P, double* X, double* Cout, double* diis_err, double* red, int nb,
                                    int mixed, void* stream) {
  if (nb < 1 || nb > PYSEQM_MAX_RED_BLOCKS) return PYSEQM_EINVAL;
  int rc = mixed ? pyseqm_gemm_bf16x3(L, idx, nact, nmax, F, P, X, stream)
                 : pyseqm_dgemm_sym(L, idx, nact, nmax, L->mol_mat_off, nullptr, F, P, X, 0, stream);
  if (rc != PYSEQM_OK || nact == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  k_antisym<<<dim3(nb, nact), kEl, 0, s>>>(*L, idx, nb, X, Cout, red);
  k_antisym_final<<<nact, 64, 0, s>>>(idx, nact, nb, red, diis_err);
  return status();
}

extern "C" int pyseqm_diis_dots(const pyseqm_layout_t* L, const int* idx, int nact, const double* ring,
                                long long ring_stride, int slot, int count, int nfock, double* Bmat, double* red,
                                int nb, void* stream) {
  if (!L || nact < 0 || count <= 0 || count > nfock || nfock > 16 || slot < 0 || slot >= nfock || nb < 1 ||
      nb > PYSEQM_MAX_RED_BLOCKS)
    return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_diis_dots<<<dim3(nb, nact), kEl, 0, s>>>(*L, idx, nb, ring, (size_t)ring_stride, slot, count, red);
  k_diis_dots_final<<<nact * count, 64, 0, s>>>(idx, nact, nb, slot, count, nfock, red,
                                                                         Bmat);
  return status();
}

extern "C" int pyseqm_diis_solve(const pyseqm_layout_t* L, const int* idx, int nact, int nfock, int count, int slot,
                                 const double* Bmat, double* coef, unsigned char* reset, void* stream) {
  if (!L || nact < 0 || count < 1 || count + 1 > kDiisMax || count > nfock || slot < 0 || slot >= nfock)
    return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  k_diis_solve<<<nact, 64, 0, (hipStream_t)stream>>>(idx, nfock, count, slot, Bmat, coef, reset);
  return status();
}

extern "C" int pyseqm_diis_combine(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* ring,
                                   long long ring_stride, int count, const double* coef, int cstride, double* F,
                                   void* stream) {
  if (!L || nact < 0 || count <= 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const int nb = tile_pair_blocks(nmax);
  k_diis_combine<<<dim3(nb, nact), kEl, 0, (hipStream_t)stream>>>(*L, idx, nb, ring, (size_t)ring_stride, count, coef,
                                                                   cstride, F);
  return status();
}
