// FP64 density step by diagonalisation, one workgroup per molecule (blocks up to 192 orbitals):
// P = 2 C_occ C_occ^T of the packed Fock block (sym_eig_trunc, seqm/seqm_functions/diag.py:110-241,
// as the SCF's make_Pnew uses it, scf_loop.py:55-103), without padding the block to a batch size.
//
//   1. Householder reduction of the block to tridiagonal form T = Q^T F Q (dsytd2 order), with the
//      matrix in registers: a GR x 16 grid of threads, thread (tr, tc) holding the RR x RC elements
//      (tr + GR a, tc + 16 b); row sums of the symmetric product reduce over the 16 lanes of a row
//      group (one 16-lane DPP row) by DPP moves, and only vectors pass through LDS.  The
//      reflectors go to the global workspace.  4 x 16 threads with 12 x 3 slots for blocks of up
//      to 48 orbitals, 16 x 16 with 6 x 6 up to 96, 32 x 16 with 6 x 12 up to 192 (64, 256, 512).
//   2. The nocc lowest eigenvalues of T by multisection on Sturm counts: a group of lanes per
//      eigenvalue evaluates the count at evenly spaced points of its bracket each round.
//   3. Their eigenvectors by inverse iteration (tridiagonal LU with partial pivoting, three solves
//      from a fixed pseudo-random start), one thread per vector, the vectors in LDS.
//   4. Vectors of eigenvalues closer than 1e-5 |T| (a cluster) orthogonalised against each other
//      by two Gram-Schmidt passes, every vector normalised.  Inverse iteration leaves vectors of
//      eigenvalues further apart orthogonal to eps |T| / gap < 3e-11; a degenerate eigenvalue's
//      vectors come from different start vectors, so they span its subspace and the passes make
//      them orthonormal inside it (P depends on the occupied subspace only).
//   5. Back-transformation C = Q Z in LDS (groups of lanes per column of Z), then P = 2 C C^T.
// Every reduction runs in a fixed order: reruns are bitwise identical.
#include "common.h"

using namespace pyseqm;

namespace {

constexpr int kEighMaxN = 192;      // largest block this solver takes
constexpr int kLdsDoubles = 20480;  // 160 KiB of LDS
constexpr int kRed = 128;           // doubles of reduction scratch (two halves of 64)

// workspace per listed molecule, in doubles, at slot * 4 * nmax^2:
//   [0, nmax^2)          reflectors: row i holds v_i (v_i[i+1] = 1) and tau_i at [i][i]
//   [nmax^2, 4 nmax^2)   the U factor of each inverse-iteration LU (reciprocal diagonal and two
//                        superdiagonals), element k of vector j at k * nmax + j
// LDS: eigenvectors n x nocc, then d, e, eigenvalues (n each), two reflector buffers and p (S each,
// S = the variant's 48 / 96 / 192 register-grid rows, zero-padded past n), kRed of scratch.  The
// variant, and so S, follows the launch's largest block nmax.
__host__ __device__ inline int grid_slots(int nmax) { return nmax <= 48 ? 48 : (nmax <= 96 ? 96 : 192); }
__host__ __device__ inline int lds_doubles(int nmax, int nocc) {
  return nmax * nocc + 3 * nmax + 3 * grid_slots(nmax) + kRed;
}

template <int NT>
__device__ inline double eig_block_max(double v, double* red) {
  constexpr int NW = NT / 64;
  for (int o = 1; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o));
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int k = 1; k < NW; ++k) s = fmax(s, red[k]);
  return s;
}

template <int NT>
__device__ inline double eig_block_min(double v, double* red) {
  constexpr int NW = NT / 64;
  for (int o = 1; o < 64; o <<= 1) v = fmin(v, __shfl_xor(v, o));
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int k = 1; k < NW; ++k) s = fmin(s, red[k]);
  return s;
}

// 1 / q to full FP64 accuracy without the scaled division sequence: v_rcp_f64 and two Newton
// steps (q is kept away from zero and overflow by the caller's pivmin)
__device__ inline double rcp_f64(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  r = fma(fma(-q, r, 1.0), r, r);
  return r;
}

// number of eigenvalues of the tridiagonal (d, e) below x: the negative pivots q_k of the LDL^T
// factorisation of T - x I (e2[k] = e[k]^2), counted as sign changes of the Sturm sequence
// p_0 = 1, p_k = (d_k - x) p_{k-1} - e2_{k-1} p_{k-2} (q_k = p_k / p_{k-1}), whose dependent chain
// is one FMA per step instead of a reciprocal.  |q_k| < pivmin is floored to q_k = -pivmin, as in
// the ratio form: |p_k| < pivmin |p_{k-1}| -> p_k = -pivmin p_{k-1}.  Every 8 steps both terms are
// scaled by the same power of two (exact; signs and ratios unchanged), which keeps them in range.
__device__ inline void sturm_step(double dk, double e2k, double x, double pivmin, double& p0, double& p1, int& cnt) {
  double p2 = fma(dk - x, p1, -(e2k * p0));
  if (fabs(p2) < pivmin * fabs(p1)) p2 = -pivmin * p1;
  cnt += (__double_as_longlong(p2) ^ __double_as_longlong(p1)) < 0;  // sign bits differ
  p0 = p1;
  p1 = p2;
}
__device__ inline int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double p0 = 1.0, p1 = d[0] - x;
  if (fabs(p1) < pivmin) p1 = -pivmin;
  int cnt = p1 < 0.0;
  int k = 1;
  for (; k + 8 <= n; k += 8) {
#pragma unroll
    for (int t = 0; t < 8; ++t) sturm_step(d[k + t], e2[k + t - 1], x, pivmin, p0, p1, cnt);
    int ex;
    frexp(fmax(fabs(p0), fabs(p1)), &ex);
    p0 = ldexp(p0, -ex);
    p1 = ldexp(p1, -ex);
  }
  for (; k < n; ++k) sturm_step(d[k], e2[k - 1], x, pivmin, p0, p1, cnt);
  return cnt;
}

// One DPP lane permutation of a double (two 32-bit DPP moves; VALU, no LDS round trip).
// Every control used here (quad permutations, row mirrors) names a source lane inside the row for
// every lane, so no lane keeps an old value and none needs initialising.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 16 lanes of a DPP row, in every lane: partners at xor 1, xor 2 (quad
// permutations), then the other quad (half-row mirror) and the other half-row (row mirror).
// Additions are commutative, so every lane ends with the same bits.
__device__ inline double row16_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return v;
}

// Sum of 32 values held by lanes 0..31 (and again by lanes 32..63): each 16-lane row sums by DPP,
// then the two rows of each half add (a + b in one lane, b + a in its partner: the same bits).
__device__ inline double sum32(double v) {
  v = row16_sum(v);
  return v + __shfl_xor(v, 16);
}

// Sum / max / min over aligned groups of t lanes (t a power of two, the same in every lane):
// DPP within a 16-lane row, shuffles across rows.
__device__ inline double group_sum_t(double v, int t) {
  if (t >= 2) v += dpp_f64<0xB1>(v);
  if (t >= 4) v += dpp_f64<0x4E>(v);
  if (t >= 8) v += dpp_f64<0x141>(v);
  if (t >= 16) v += dpp_f64<0x140>(v);
  if (t >= 32) v += __shfl_xor(v, 16);
  if (t >= 64) v += __shfl_xor(v, 32);
  return v;
}
__device__ inline double group_max_t(double v, int t) {
  if (t >= 2) v = fmax(v, dpp_f64<0xB1>(v));
  if (t >= 4) v = fmax(v, dpp_f64<0x4E>(v));
  if (t >= 8) v = fmax(v, dpp_f64<0x141>(v));
  if (t >= 16) v = fmax(v, dpp_f64<0x140>(v));
  if (t >= 32) v = fmax(v, __shfl_xor(v, 16));
  if (t >= 64) v = fmax(v, __shfl_xor(v, 32));
  return v;
}
__device__ inline double group_min_t(double v, int t) {
  if (t >= 2) v = fmin(v, dpp_f64<0xB1>(v));
  if (t >= 4) v = fmin(v, dpp_f64<0x4E>(v));
  if (t >= 8) v = fmin(v, dpp_f64<0x141>(v));
  if (t >= 16) v = fmin(v, dpp_f64<0x140>(v));
  if (t >= 32) v = fmin(v, __shfl_xor(v, 16));
  if (t >= 64) v = fmin(v, __shfl_xor(v, 32));
  return v;
}

// deterministic pseudo-random start vector element in (-1, 1)
__device__ inline double start_value(int j, int k) {
  unsigned h = (unsigned)j * 2654435761u + (unsigned)k * 40503u + 12345u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  return (double)h * (2.0 / 4294967296.0) - 1.0;
}

// C = Q Z on the register block (thread (tr, tc) holds C[tc + 16 b][tr + GR c]): the reflectors
// last to first, C -= v (tau v^T C).  Only the first CMAX column slots hold vectors (a compile-time
// count, so the work follows nv).  Every row slot takes part in every step: v is zero at rows <= i
// and beyond the block, so those products and updates are exact no-ops, and the loops stay free
// of per-slot selects; the CMAX column chains run interleaved.  The next reflector is read from
// the workspace while the current one is applied.  C (in LDS, n x nv, over Z) is written after a
// barrier that follows every thread's read of Z.
template <int GR, int RR, int RC, int CMAX>
__device__ __forceinline__ void back_transform(double (&Ar)[RR][RC], const double* __restrict__ HH, int nmax, int n,
                                               int nv, int tr, int tc, double* C) {
  double vcur[RC], vnext[RC];
  double tcur = 0.0, tnext = 0.0;
  auto load_v = [&](int i, double* v, double& t) {
    t = HH[(size_t)i * nmax + i];
#pragma unroll
    for (int b = 0; b < RC; ++b) {
      const int k = tc + 16 * b;
      v[b] = (k > i && k < n) ? HH[(size_t)i * nmax + k] : 0.0;
    }
  };
  if (n >= 3) load_v(n - 3, vcur, tcur);
  for (int i = n - 3; i >= 0; --i) {
    if (i >= 1) load_v(i - 1, vnext, tnext);
    if (tcur != 0.0) {  // the same in every thread
      double sd[CMAX];
#pragma unroll
      for (int c = 0; c < CMAX; ++c) sd[c] = 0.0;
#pragma unroll
      for (int b = 0; b < RC; ++b)
#pragma unroll
        for (int c = 0; c < CMAX; ++c) sd[c] = fma(Ar[c][b], vcur[b], sd[c]);
#pragma unroll
      for (int c = 0; c < CMAX; ++c) sd[c] = tcur * row16_sum(sd[c]);
#pragma unroll
      for (int b = 0; b < RC; ++b)
#pragma unroll
        for (int c = 0; c < CMAX; ++c) Ar[c][b] = fma(-vcur[b], sd[c], Ar[c][b]);
    }
#pragma unroll
    for (int b = 0; b < RC; ++b) vcur[b] = vnext[b];
    tcur = tnext;
  }
  __syncthreads();  // every thread has read Z
#pragma unroll
  for (int c = 0; c < CMAX; ++c)
#pragma unroll
    for (int b = 0; b < RC; ++b) {
      const int k = tc + 16 * b, a = tr + GR * c;
      if (k < n && a < nv) C[k * nv + a] = Ar[c][b];
    }
}

// EVEC = false: the density P = 2 C_occ C_occ^T of the nocc lowest vectors.  EVEC = true: every
// eigenpair (nv = n vectors), written as Qout (packed layout, column j = eigenvector j) and evals
// (orb_off[m] + j, the Rayleigh quotients of the final vectors: exact to rounding whatever the
// multisection bracket width).
// CM: column slots of the back-transformation that hold vectors (GR CM >= nv of every listed block,
// chosen per launch from the launch's largest nv).
template <int GR, int GC, int RR, int RC, bool EVEC, int CM>
__global__ __launch_bounds__(GR* GC) void k_eigh_density(pyseqm_layout_t L, const int* __restrict__ idx,
                                                       const double* __restrict__ F, double* __restrict__ P,
                                                       double* __restrict__ ws, int nmax,
                                                       int* __restrict__ fail, const long long* __restrict__ orb_off,
                                                       double* __restrict__ evals) {
  constexpr int NT = GR * GC;
  static_assert(GC == 16, "row groups are DPP rows");
  extern __shared__ double lds[];
  const int slot = blockIdx.x;
  const int m = idx[slot];
  const int n = L.mol_norb[m];
  const int nv = EVEC ? n : L.mol_nocc[m];  // vectors computed
  const size_t mo = L.mol_mat_off[m];
  const int tid = threadIdx.x;
  const size_t nm2 = (size_t)nmax * nmax;
  double* HH = ws + (size_t)slot * 4 * nm2;
  double* U0 = HH + nm2;
  double* U1 = HH + 2 * nm2;
  double* U2 = HH + 3 * nm2;
  double* Z = lds;  // eigenvectors, n x nv: Z[k * nv + j]
  double* d = Z + n * nv;
  double* e = d + n;
  double* lam = e + n;
  constexpr int S = GR * RR;  // = GC * RC: register-grid rows / columns
  static_assert(GR * RR == GC * RC, "square register grid");
  double* xs0 = lam + n;  // reflector columns of even / odd steps, zero at indices <= i + 1 and >= n
  double* xs1 = xs0 + S;
  double* pp = xs1 + S;
  double* red = pp + S;  // kRed
  double* red1 = red;  // GR <= 64 partials each
  double* red2 = red + kRed / 2;
  const double* Fm = F + mo;

  // ---- 1. Householder tridiagonalisation (dsytd2, lower) in registers, two barriers per step.
  // Thread (tr, tc) of the GR x GC grid holds A[tr + GR a][tc + GC b], a < RR, b < RC.
  // Step i: x = A[i+1:, i] is in LDS with partial sums of |x[1:]|^2 (written by the owners of
  // column i, tc == i mod GC, during the previous step's update); every thread forms tau, beta
  // and v = (1, x[1:] / (alpha - beta)) at its columns;
  // p = tau A v (row sums over the lanes of each row group, rows <= i masked by v = 0), with
  // partial sums of p.v; then A -= v w^T + w v^T, w = p - (tau / 2)(p.v) v, on every slot (v = w
  // = 0 at indices <= i leave the finished rows and columns alone).
  const int tr = tid / GC, tc = tid - (tid / GC) * GC;
  double Ar[RR][RC];
#pragma unroll
  for (int a = 0; a < RR; ++a)
#pragma unroll
    for (int b = 0; b < RC; ++b) {
      const int r = tr + GR * a, c = tc + GC * b;
      Ar[a][b] = (r < n && c < n) ? Fm[(size_t)r * n + c] : 0.0;
    }
  // the reflector buffers and p start at zero (their entries past n stay zero), then column 0
  // (the owners of column 0, tc == 0, slot b = 0); later columns are written by the rank-2
  // update of the step before, which computes them
  for (int t = tid; t < 3 * S; t += NT) xs0[t] = 0.0;
  __syncthreads();
  if (tc == 0) {
    double part = 0.0;
#pragma unroll
    for (int a = 0; a < RR; ++a) {
      const int r = tr + GR * a;
      const double val = Ar[a][0];
      if (r == 0) {
        d[0] = val;
      } else if (r < n) {
        xs0[r] = val;
        if (r > 1) part += val * val;
      }
    }
    red1[tr] = part;
  }
  const int lane32 = tid & 31;
  for (int i = 0; i + 2 < n; ++i) {
    const double* x = (i & 1) ? xs1 : xs0;
    double* xnext = (i & 1) ? xs0 : xs1;
    __syncthreads();
    // every LDS read of the step's first half is issued here, unconditionally, so that they travel
    // together instead of one round trip per slot (x holds S entries, zero past the live range)
    double xc[RC], xr[RR];
#pragma unroll
    for (int b = 0; b < RC; ++b) xc[b] = x[tc + GC * b];
#pragma unroll
    for (int a = 0; a < RR; ++a) xr[a] = x[tr + GR * a];
    const double alpha = x[i + 1];
    // |x[1:]|^2 from the GR partials: lane k of every wave takes partial k mod 32, then a fixed
    // tree (16-lane DPP row, then the row pair): the same bits in every thread
    const double r1 = red1[lane32];  // red1 holds kRed / 2 = 64 doubles: in bounds
    const double xn2 = sum32(lane32 < GR ? r1 : 0.0);
    double tau = 0.0, beta = alpha, scal = 0.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      const double am = alpha - beta;
      scal = rcp_f64(am);
      tau = -am * rcp_f64(beta);
    }
    // tau = 0 (column already zero) runs the same steps with p = w = 0: A unchanged
    if (tid == 0) {
      e[i] = beta;
      HH[(size_t)i * nmax + i] = tau;
    }
    // v at this thread's columns and rows: 1 at i + 1, else x scal (x is zero at indices <= i and
    // >= n, so no range test is needed)
    double vc[RC], vr[RR];
#pragma unroll
    for (int b = 0; b < RC; ++b) vc[b] = (tc + GC * b == i + 1) ? 1.0 : xc[b] * scal;
#pragma unroll
    for (int a = 0; a < RR; ++a) vr[a] = (tr + GR * a == i + 1) ? 1.0 : xr[a] * scal;
    if (tr == 0) {
#pragma unroll
      for (int b = 0; b < RC; ++b) {
        const int c = tc + GC * b;
        if (c > i && c < n) HH[(size_t)i * nmax + c] = vc[b];
      }
    }
    // p = tau A v.  Column slots whose columns are all finished (v = 0) or all beyond the block are
    // skipped (a uniform test); every row of a live slot is summed: finished rows and rows beyond
    // the block give values that are never used (p is kept at live rows only).
    double s[RR];
#pragma unroll
    for (int a = 0; a < RR; ++a) s[a] = 0.0;
#pragma unroll
    for (int b = 0; b < RC; ++b) {
      if (GC * b + GC - 1 <= i || GC * b >= n) continue;
#pragma unroll
      for (int a = 0; a < RR; ++a) s[a] = fma(Ar[a][b], vc[b], s[a]);
    }
    // p is written at every row (zero off the live range, finished slots included), so that w
    // below needs no range test
    double q = 0.0;
#pragma unroll
    for (int a = 0; a < RR; ++a) {
      double pa = 0.0;
      if (GR * a + GR - 1 > i) {  // some row of the slot still live (uniform)
        pa = tau * row16_sum(s[a]);
        q = fma(pa, vr[a], q);  // vr = 0 off the live rows
      }
      const int r = tr + GR * a;
      if (tc == 0) pp[r] = (r > i && r < n) ? pa : 0.0;
    }
    if (tc == 0) red2[tr] = q;
    __syncthreads();
    // p at this thread's rows and at the first chunk of its columns are read with the partials;
    // each later chunk's columns are read while the chunk before it is applied
    constexpr int kUpdChunk = 4;
    constexpr int kChunks = (RC + kUpdChunk - 1) / kUpdChunk;
    double wr[RR], wc[2][kUpdChunk];
#pragma unroll
    for (int a = 0; a < RR; ++a) wr[a] = pp[tr + GR * a];
#pragma unroll
    for (int j = 0; j < kUpdChunk; ++j)
      if (j < RC) wc[0][j] = pp[tc + GC * j];
    const double r2 = red2[lane32];
    const double pv = sum32(lane32 < GR ? r2 : 0.0);
    const double a2 = -0.5 * tau * pv;
#pragma unroll
    for (int a = 0; a < RR; ++a) wr[a] = fma(a2, vr[a], wr[a]);  // p = v = 0 off the live rows
    // A -= v w^T + w v^T on the live column slots, every row (v = w = 0 at finished rows and rows
    // beyond the block leave them exactly unchanged).  Column i + 1 (slot bn, a uniform index) is
    // then copied out and handed to the next step by its owners (tc == tn): the reflector input x,
    // its partial squared norm and the diagonal element.
    const int bn = (i + 1) / GC, tn = (i + 1) - bn * GC;
    double coln[RR];
#pragma unroll
    for (int a = 0; a < RR; ++a) coln[a] = 0.0;
#pragma unroll
    for (int ch = 0; ch < kChunks; ++ch) {
      const int b0 = ch * kUpdChunk;
      if (ch + 1 < kChunks) {
#pragma unroll
        for (int j = 0; j < kUpdChunk; ++j)
          if (b0 + kUpdChunk + j < RC) wc[(ch + 1) & 1][j] = pp[tc + GC * (b0 + kUpdChunk + j)];
      }
      if (GC * min(b0 + kUpdChunk, RC) - 1 <= i || GC * b0 >= n) continue;  // whole chunk idle (uniform)
#pragma unroll
      for (int j = 0; j < kUpdChunk; ++j) {
        const int b = b0 + j;
        if (b >= RC) break;
        if (GC * b + GC - 1 <= i || GC * b >= n) continue;
        const double vb = vc[b];
        const double wb = fma(a2, vb, wc[ch & 1][j]);  // p = v = 0 off the live columns
#pragma unroll
        for (int a = 0; a < RR; ++a) Ar[a][b] = fma(-vr[a], wb, fma(-wr[a], vb, Ar[a][b]));
      }
    }
#pragma unroll
    for (int b = 0; b < RC; ++b)
      if (b == bn) {  // uniform
#pragma unroll
        for (int a = 0; a < RR; ++a) coln[a] = Ar[a][b];
      }
    double partn = 0.0;
    if (tc == tn) {
#pragma unroll
      for (int a = 0; a < RR; ++a) {
        const int r = tr + GR * a;
        if (r == i + 1) d[i + 1] = coln[a];
        xnext[r] = (r > i + 1 && r < n) ? coln[a] : 0.0;  // every row: zero off the live range
        if (r > i + 2 && r < n) partn = fma(coln[a], coln[a], partn);
      }
    }
    if (tc == tn) red1[tr] = partn;
  }
  // the trailing 2 x 2 block (or the 1 x 1 of a one-orbital block)
#pragma unroll
  for (int a = 0; a < RR; ++a)
#pragma unroll
    for (int b = 0; b < RC; ++b) {
      const int r = tr + GR * a, c = tc + GC * b;
      if (r == n - 2 && c == n - 2) d[n - 2] = Ar[a][b];
      if (r == n - 1 && c == n - 2) e[n - 2] = Ar[a][b];
      if (r == n - 1 && c == n - 1) d[n - 1] = Ar[a][b];
    }
  __syncthreads();

  // ---- 2. the nv lowest eigenvalues: multisection, a group of tb lanes per eigenvalue
  double lo_p = 1.0e300, hi_p = -1.0e300, e2max = 0.0;
  for (int k = tid; k < n; k += NT) {
    const double el = k > 0 ? fabs(e[k - 1]) : 0.0, er = k + 1 < n ? fabs(e[k]) : 0.0;
    lo_p = fmin(lo_p, d[k] - el - er);
    hi_p = fmax(hi_p, d[k] + el + er);
    if (k + 1 < n) {
      pp[k] = e[k] * e[k];  // e^2 for the Sturm counts (pp is free after the reduction)
      e2max = fmax(e2max, pp[k]);
    }
  }
  double glo = eig_block_min<NT>(lo_p, red);
  double ghi = eig_block_max<NT>(hi_p, red);
  e2max = eig_block_max<NT>(e2max, red);
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  const double pivmin = 1.0e-300 * fmax(1.0, e2max);
  glo -= 2.2e-16 * tnorm * 4.0 + 2.0 * pivmin;
  ghi += 2.2e-16 * tnorm * 4.0 + 2.0 * pivmin;
  // Bracket width at which an eigenvalue is final: 1e-11 |T| (or the rounding floor).  Inverse
  // iteration only needs the shift that close: with eigenvalues further apart than the cluster
  // threshold 1e-5 |T| each solve shrinks the other components by 1e-6 or more, and the vectors
  // of a cluster are orthogonalised as a set.
  const double btol = 1.0e-11 * tnorm;
  {
    int tb = 64;
    while (tb > 1 && nv * tb > NT) tb >>= 1;
    const int ngrp = NT / tb, gi = tid / tb, gl = tid % tb;
    for (int j0 = 0; j0 < nv; j0 += ngrp) {
      const int j = j0 + gi;
      if (j < nv) {  // uniform over the lane group
        double lo = glo, hi = ghi;
        for (int round = 0; round < 64; ++round) {
          if (!(hi - lo > fmax(4.4e-16 * fmax(fabs(lo), fabs(hi)), btol) + 2.0 * pivmin)) break;
          const double x = lo + (hi - lo) * (double)(gl + 1) / (double)(tb + 1);
          const int cnt = (x > lo && x < hi) ? sturm_count(d, pp, n, x, pivmin) : -1;
          const double nlo = group_max_t(cnt >= 0 && cnt <= j ? x : lo, tb);
          const double nhi = group_min_t(cnt > j ? x : hi, tb);
          if (nlo == lo && nhi == hi) break;  // no representable interior point left
          lo = nlo;
          hi = nhi;
        }
        if (gl == 0) lam[j] = 0.5 * (lo + hi);
      }
    }
  }
  __syncthreads();

  // ---- 3. inverse iteration, one thread per eigenvector; vector j in LDS at Z[k * nv + j]
  const double tiny = 2.2e-16 * fmax(tnorm, 1.0e-300);
  for (int j = tid; j < nv; j += NT) {
    for (int k = 0; k < n; ++k) Z[k * nv + j] = start_value(j, k);
    const double sig = lam[j];
    for (int rep = 0; rep < 3; ++rep) {
      // LU of T - sig I with partial pivoting, the right-hand side eliminated alongside; the
      // running row and right-hand side stay in registers, U goes to the workspace.  Each step
      // keeps one of two rows as the pivot row (its diagonal and sub-diagonal entries compared)
      // and eliminates the other, both cases as one select-based update (lanes differ in their
      // choices); the next step's d, e and right-hand side are read from LDS one step ahead.
      double diag = d[0] - sig, sup = n > 1 ? e[0] : 0.0;
      double yk = Z[j];
      double sub_n = n > 1 ? e[0] : 0.0, dn_n = n > 1 ? d[1] : 0.0, ns_n = n > 2 ? e[1] : 0.0;
      double y_n = n > 1 ? Z[nv + j] : 0.0;
      for (int k = 0; k + 1 < n; ++k) {
        const double sub = sub_n, nd = dn_n - sig, ns = ns_n, yk1 = y_n;
        if (k + 2 < n) {
          sub_n = ns_n;
          dn_n = d[k + 2];
          ns_n = k + 3 < n ? e[k + 2] : 0.0;
          y_n = Z[(k + 2) * nv + j];
        }
        const size_t o = (size_t)k * nmax + j;
        const bool keep = fabs(diag) >= fabs(sub);  // no row exchange
        const double piv = keep ? (fabs(diag) < tiny ? copysign(tiny, diag) : diag) : sub;
        const double rp = rcp_f64(piv);
        const double mlt = (keep ? sub : diag) * rp;
        const double p1 = keep ? sup : nd, p2 = keep ? 0.0 : ns, py = keep ? yk : yk1;
        const double o1 = keep ? nd : sup, o2 = keep ? ns : 0.0, oy = keep ? yk1 : yk;
        U0[o] = rp;
        U1[o] = p1;
        U2[o] = p2;
        Z[k * nv + j] = py;
        yk = fma(-mlt, py, oy);
        diag = fma(-mlt, p1, o1);
        sup = fma(-mlt, p2, o2);
      }
      const double rlast = 1.0 / (fabs(diag) < tiny ? copysign(tiny, diag) : diag);
      double x1 = yk * rlast, x2 = 0.0, xmax = fabs(x1);
      Z[(n - 1) * nv + j] = x1;
      // back substitution in blocks of kPF rows: the U factors of the next block are loaded while
      // this block is solved (the loads come from L2; one row ahead left each row waiting on one)
      constexpr int kPF = 8;
      double c0[kPF], c1[kPF], c2[kPF];
#pragma unroll
      for (int t = 0; t < kPF; ++t) {
        const int kk = n - 2 - t;
        const size_t o = (size_t)(kk >= 0 ? kk : 0) * nmax + j;
        c0[t] = U0[o];
        c1[t] = U1[o];
        c2[t] = U2[o];
      }
      for (int k0 = n - 2; k0 >= 0; k0 -= kPF) {
        double n0[kPF], n1[kPF], n2[kPF];
#pragma unroll
        for (int t = 0; t < kPF; ++t) {
          const int kk = k0 - kPF - t;
          const size_t o = (size_t)(kk >= 0 ? kk : 0) * nmax + j;
          n0[t] = U0[o];
          n1[t] = U1[o];
          n2[t] = U2[o];
        }
#pragma unroll
        for (int t = 0; t < kPF; ++t) {
          const int k = k0 - t;
          if (k >= 0) {
            const double x = (Z[k * nv + j] - c1[t] * x1 - c2[t] * x2) * c0[t];
            Z[k * nv + j] = x;
            xmax = fmax(xmax, fabs(x));
            x2 = x1;
            x1 = x;
          }
        }
#pragma unroll
        for (int t = 0; t < kPF; ++t) {
          c0[t] = n0[t];
          c1[t] = n1[t];
          c2[t] = n2[t];
        }
      }
      const double sc = xmax > 0.0 ? 1.0 / xmax : 1.0;
      for (int k = 0; k < n; ++k) Z[k * nv + j] *= sc;
    }
  }
  __syncthreads();

  // ---- 4. clusters (eigenvalues closer than ctol) orthogonalised by the thread of their first
  //         member, two Gram-Schmidt passes; every vector normalised
  const double ctol = 1.0e-5 * fmax(tnorm, 1.0e-300);
  for (int j = tid; j < nv; j += NT) {
    if (j > 0 && lam[j] - lam[j - 1] <= ctol) continue;  // inside a cluster: its first thread works
    int jend = j + 1;
    while (jend < nv && lam[jend] - lam[jend - 1] <= ctol) ++jend;
    for (int q = j; q < jend; ++q) {
      for (int pass = 0; pass < 2; ++pass)
        for (int p = j; p < q; ++p) {
          double dot = 0.0;
          for (int k = 0; k < n; ++k) dot += Z[k * nv + p] * Z[k * nv + q];
          for (int k = 0; k < n; ++k) Z[k * nv + q] -= dot * Z[k * nv + p];
        }
      double nr = 0.0;
      for (int k = 0; k < n; ++k) nr += Z[k * nv + q] * Z[k * nv + q];
      if (!(nr > 1.0e-200)) {  // a vector lost to its cluster's span: flagged, left unscaled
        atomicAdd(fail, 1);
        continue;
      }
      const double sc = 1.0 / sqrt(nr);
      for (int k = 0; k < n; ++k) Z[k * nv + q] *= sc;
    }
  }
  __syncthreads();

  if (EVEC) {  // Rayleigh quotients z^T T z of the normalised vectors (ascending to rounding)
    for (int j = tid; j < nv; j += NT) {
      double rq = d[0] * Z[j] * Z[j];
      for (int k = 1; k < n; ++k) {
        const double zk = Z[k * nv + j], zk1 = Z[(k - 1) * nv + j];
        rq += d[k] * zk * zk + 2.0 * e[k - 1] * zk * zk1;
      }
      lam[j] = rq;
    }
    __syncthreads();
  }
  // ---- 5. C = Q Z in registers: thread (tr, tc) holds C[tc + 16 b][tr + GR c] (b < RC rows
  //         over the lanes of its DPP row, c < RR columns), so each column's product v^T C reduces
  //         inside one DPP row and the reflectors (last to first) need no barrier; each thread
  //         reads its RC elements of the next reflector from the workspace while applying one.
  double* C = Z;
  {
#pragma unroll
    for (int c = 0; c < RR; ++c)
#pragma unroll
      for (int b = 0; b < RC; ++b) {
        const int k = tc + 16 * b, a = tr + GR * c;
        Ar[c][b] = (k < n && a < nv) ? Z[k * nv + a] : 0.0;
      }
    back_transform<GR, RR, RC, CM>(Ar, HH, nmax, n, nv, tr, tc, C);
    __syncthreads();
  }
  if (EVEC) {
    double* Qm = P + mo;
    for (int t = tid; t < n * n; t += NT) Qm[t] = C[t];  // C is n x n, row k = orbital, column j = vector
    for (int j = tid; j < n; j += NT) evals[orb_off[m] + j] = lam[j];
    return;
  }
  // ---- P = 2 C C^T, 4 x 4 output blocks per thread over the upper triangle of blocks (tasks
  //      enumerated over the triangle only), both triangles written; every element sums over the
  //      vectors in order, and the two copies of a diagonal block's element are the same bits
  double* Pm = P + mo;
  const int nb4 = (n + 3) / 4;
  const int ntask = nb4 * (nb4 + 1) / 2;
  for (int t = tid; t < ntask; t += NT) {
    int bi = 0, rem = t;
    while (rem >= nb4 - bi) {
      rem -= nb4 - bi;
      ++bi;
    }
    const int bj = bi + rem;
    int ro[4], co[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ro[q] = min(4 * bi + q, n - 1) * nv;
      co[q] = min(4 * bj + q, n - 1) * nv;
    }
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
    for (int a = 0; a < nv; ++a) {
      double x[4], y[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        x[q] = C[ro[q] + a];
        y[q] = C[co[q] + a];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] += x[p] * y[q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = 4 * bi + p, c = 4 * bj + q;
        if (r < n && c < n) {
          Pm[(size_t)r * n + c] = 2.0 * acc[p][q];
          Pm[(size_t)c * n + r] = 2.0 * acc[p][q];
        }
      }
  }
}

}  // namespace

extern "C" long long pyseqm_eigh_density_workspace(int nact, int nmax) {
  if (nact < 0 || nmax < 0) return -1;
  return 4LL * nact * nmax * nmax;
}

// thread grid GR x 16 with RR x RC register slots per thread: GR RR and 16 RC cover nmax
// (the back-transformation's column slots: half of them when every listed block's vector count
// nvmax fits, which the density step's nocc usually does)
template <int GR, int RR, int RC, bool EVEC>
void launch_variant(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int nvmax, const double* F,
                    double* out, double* ws, int* fail, const long long* orb_off, double* evals, size_t lds,
                    hipStream_t s) {
  if (nvmax <= GR * (RR / 2))
    k_eigh_density<GR, 16, RR, RC, EVEC, RR / 2><<<nact, GR * 16, lds, s>>>(*L, idx, F, out, ws, nmax, fail, orb_off,
                                                                         evals);
  else
    k_eigh_density<GR, 16, RR, RC, EVEC, RR><<<nact, GR * 16, lds, s>>>(*L, idx, F, out, ws, nmax, fail, orb_off,
                                                                      evals);
}

template <bool EVEC>
void launch_eigh(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int nvmax, const double* F,
                 double* out, double* ws, int* fail, const long long* orb_off, double* evals, size_t lds,
                 hipStream_t s) {
  if (nmax <= 48)
    launch_variant<4, 12, 3, EVEC>(L, idx, nact, nmax, nvmax, F, out, ws, fail, orb_off, evals, lds, s);
  else if (nmax <= 96)
    launch_variant<16, 6, 6, EVEC>(L, idx, nact, nmax, nvmax, F, out, ws, fail, orb_off, evals, lds, s);
  else
    launch_variant<32, 6, 12, EVEC>(L, idx, nact, nmax, nvmax, F, out, ws, fail, orb_off, evals, lds, s);
}

extern "C" int pyseqm_eigh_density_lds_ok(int norb, int nocc) {
  if (norb < 1 || norb > kEighMaxN || nocc < 0 || nocc > norb) return 0;
  return lds_doubles(norb, nocc) <= kLdsDoubles;
}

extern "C" int pyseqm_eigh_density(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int nocc_max,
                                   const double* F, double* P, double* ws, int* fail, void* stream) {
  if (!L || nact < 0 || nmax < 1 || nmax > kEighMaxN || nocc_max < 0 || nocc_max > nmax || !fail)
    return PYSEQM_EINVAL;
  if (!pyseqm_eigh_density_lds_ok(nmax, nocc_max)) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const size_t lds = (size_t)lds_doubles(nmax, nocc_max) * sizeof(double);
  hipStream_t s = (hipStream_t)stream;
  launch_eigh<false>(L, idx, nact, nmax, nocc_max, F, P, ws, fail, nullptr, nullptr, lds, s);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_eigh_vectors(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* F,
                                   const long long* orb_off, double* evals, double* Q, double* ws, int* fail,
                                   void* stream) {
  if (!L || nact < 0 || nmax < 1 || nmax > kEighMaxN || !fail || !orb_off || !evals) return PYSEQM_EINVAL;
  if (!pyseqm_eigh_density_lds_ok(nmax, nmax)) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const size_t lds = (size_t)lds_doubles(nmax, nmax) * sizeof(double);
  launch_eigh<true>(L, idx, nact, nmax, nmax, F, Q, ws, fail, orb_off, evals, lds, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}
