// Thermal (Fermi) density and the Krylov-subspace (KSA) XL-BOMD kernel on the packed batch.
//
//   * Fermi occupations f = 1 / (1 + exp((e - mu) / kT)) of every eigenpair, the chemical potential
//     by Newton steps from the HOMO/LUMO midpoint with the reference's batch-wide stop, the density
//     D = 2 Q diag(f) Q^T and the entropy (fermi_q.py:8-72).
//   * FP64 products of n x n blocks (C = op(A) op(B)) for the eigenbasis transforms of the
//     canonical density-matrix perturbation theory, its recursion in the eigenbasis and the
//     chemical-potential shift (canon_dm_prt.py:6-39).
//   * Per-molecule Frobenius products, block updates and small dense solves of the rank-m
//     Krylov fit (seqm/dynamics/xlbomd.py:243-339).
// Every block is a molecule's packed n x n matrix at mol_mat_off; eigenvalues and occupations
// are at orb_off[m] + j.  Reductions run in a fixed order: reruns are bitwise identical.
#include "common.h"

using namespace pyseqm;

namespace {

constexpr int kFermiIter = 64;      // fermi_q.py:12
constexpr double kOccTol = 1.0e-9;  // fermi_q.py:14 (float64)
constexpr double kEntEps = 1.0e-14;
constexpr double kTiny = 1.0e-30;
constexpr int kMaxRank = 8;         // largest Krylov rank of pyseqm_small_solve

__device__ inline double sigm(double x) { return 1.0 / (1.0 + exp(-x)); }

// fixed-order sum over a 256-thread block; every thread gets the total
__device__ inline double block_sum256(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// Newton steps for mu, one wave per molecule: every step's mu and whether |nocc - sum f| <= tol
// are recorded (the batch stops at the first step at which every molecule is converged).
__global__ __launch_bounds__(64) void k_fermi_newton(pyseqm_layout_t L, const int* __restrict__ idx,
                                                     const long long* __restrict__ orb_off,
                                                     const double* __restrict__ evals, double beta,
                                                     double* __restrict__ mu_hist, unsigned char* __restrict__ conv) {
  const int slot = blockIdx.x, lane = threadIdx.x;
  const int m = idx[slot];
  const int n = L.mol_norb[m], N = L.mol_nocc[m];
  const double* e = evals + orb_off[m];
  double mu = 0.5 * (e[N - 1] + e[N]);
  for (int t = 0; t < kFermiIter; ++t) {
    double occ = 0.0, docc = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double f = sigm(-beta * (e[i] - mu));
      occ += f;
      docc += beta * f * (1.0 - f);
    }
    occ = wave_sum(occ);
    docc = fmax(wave_sum(docc), kTiny);
    if (lane == 0) {
      mu_hist[(size_t)slot * kFermiIter + t] = mu;
      conv[(size_t)slot * kFermiIter + t] = fabs((double)N - occ) <= kOccTol;
    }
    mu += ((double)N - occ) / docc;
  }
}

// The first step at which every listed molecule is converged (the last step if none is).
__global__ __launch_bounds__(256) void k_fermi_pick(int nact, const unsigned char* __restrict__ conv,
                                                    int* __restrict__ kout) {
  int K = kFermiIter - 1;
  for (int t = 0; t < kFermiIter; ++t) {
    int all = 1;
    for (int s = threadIdx.x; s < nact; s += blockDim.x) all &= conv[(size_t)s * kFermiIter + t];
    if (__syncthreads_and(all)) {
      K = t;
      break;
    }
  }
  if (threadIdx.x == 0) *kout = K;
}

// Occupations, entropy and D = 2 Q diag(f) Q^T at the chosen step's mu, one block per molecule.
__global__ __launch_bounds__(256) void k_fermi_density(pyseqm_layout_t L, const int* __restrict__ idx,
                                                       const long long* __restrict__ orb_off,
                                                       const double* __restrict__ evals,
                                                       const double* __restrict__ Q, double beta, double kB,
                                                       const double* __restrict__ mu_hist,
                                                       const int* __restrict__ kpick, double* __restrict__ occ,
                                                       double* __restrict__ mu_out, double* __restrict__ D,
                                                       double* __restrict__ S) {
  __shared__ double f[256];
  __shared__ double red[4];
  const int slot = blockIdx.x, tid = threadIdx.x;
  const int m = idx[slot];
  const int n = L.mol_norb[m];
  const size_t mo = L.mol_mat_off[m];
  const double mu = mu_hist[(size_t)slot * kFermiIter + *kpick];
  const double* e = evals + orb_off[m];
  double s = 0.0;
  for (int i = tid; i < n; i += 256) {
    const double fi = sigm(-beta * (e[i] - mu));
    f[i] = fi;
    occ[orb_off[m] + i] = fi;
    if (fi > kEntEps && 1.0 - fi > kEntEps) s += -kB * (fi * log(fi) + (1.0 - fi) * log(1.0 - fi));
  }
  s = block_sum256(s, red);  // its barriers also publish f
  if (tid == 0) {
    S[m] = s;
    mu_out[m] = mu;
  }
  const double* Qm = Q + mo;
  for (int t = tid; t < n * n; t += 256) {
    const int r = t / n, c = t - (t / n) * n;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += Qm[(size_t)r * n + j] * f[j] * Qm[(size_t)c * n + j];
    D[mo + t] = 2.0 * acc;
  }
}

// C = op(A) op(B) per listed molecule, 16 x 16 output tiles through LDS.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_block_gemm(pyseqm_layout_t L, const int* __restrict__ idx,
                                                    const double* __restrict__ A, const double* __restrict__ B,
                                                    double* __restrict__ C) {
  __shared__ double As[16][17], Bs[16][17];
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const int nt = (n + 15) / 16;
  if ((int)blockIdx.x >= nt * nt) return;  // uniform over the block
  const int ti = blockIdx.x / nt, tj = blockIdx.x - (blockIdx.x / nt) * nt;
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const size_t mo = L.mol_mat_off[m];
  const double* Am = A + mo;
  const double* Bm = B + mo;
  const int r = ti * 16 + ty, c = tj * 16 + tx;
  double acc = 0.0;
  for (int k0 = 0; k0 < n; k0 += 16) {
    const int ka = k0 + tx, kb = k0 + ty;
    As[ty][tx] = (r < n && ka < n) ? (TA ? Am[(size_t)ka * n + r] : Am[(size_t)r * n + ka]) : 0.0;
    Bs[ty][tx] = (kb < n && c < n) ? (TB ? Bm[(size_t)c * n + kb] : Bm[(size_t)kb * n + c]) : 0.0;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) acc += As[ty][kk] * Bs[kk][tx];
    __syncthreads();
  }
  if (r < n && c < n) C[mo + (size_t)r * n + c] = acc;
}

// Recursive Fermi-operator expansion of the perturbation in the eigenbasis, element by element:
// p0 = 1/2 - 2^(-2-m) beta (e - mu), P1 = -2^(-2-m) beta Y, then m steps of
// dX1 = p0_i P1 + P1 p0_j, p0 <- p0^2 / (2 (p0^2 - p0) + 1), P1 <- (dX1 + 2 (P1 - dX1) p0_j) / (2 (p0_i^2 - p0_i) + 1).
__global__ __launch_bounds__(256) void k_canon_elem(pyseqm_layout_t L, const int* __restrict__ idx,
                                                    const long long* __restrict__ orb_off,
                                                    const double* __restrict__ evals, const double* __restrict__ mu,
                                                    double beta, int steps, double* __restrict__ Y) {
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * n) return;
  const int i = t / n, j = t - (t / n) * n;
  const double* e = evals + orb_off[m];
  const double cn = ldexp(beta, -2 - steps);
  double p0i = 0.5 - cn * (e[i] - mu[m]);
  double p0j = 0.5 - cn * (e[j] - mu[m]);
  double* y = Y + L.mol_mat_off[m] + t;
  double P1 = -cn * *y;
  for (int s = 0; s < steps; ++s) {
    const double p02i = p0i * p0i, p02j = p0j * p0j;
    const double dX1 = p0i * P1 + P1 * p0j;
    const double iDi = 1.0 / (2.0 * (p02i - p0i) + 1.0);
    const double iDj = 1.0 / (2.0 * (p02j - p0j) + 1.0);
    p0i = iDi * p02i;
    p0j = iDj * p02j;
    P1 = iDi * (dX1 + 2.0 * (P1 - dX1) * p0j);
  }
  *y = P1;
}

// First-order response of the closed-shell density D = 2 sum_occ c c^T to a symmetric Fock
// perturbation, in the eigenbasis: dD'_ia = dD'_ai = 2 F'_ia / (e_i - e_a) for i occupied and a
// virtual (the nocc lowest eigenpairs), zero in the occupied-occupied and virtual-virtual blocks.
// The operator is self-adjoint, so it is also the transpose the SCF implicit backward applies
// (the eigh backward of sym_eig_trunc1 inside scf_loop.py:1779-1993).
__global__ __launch_bounds__(256) void k_response_elem(pyseqm_layout_t L, const int* __restrict__ idx,
                                                       const long long* __restrict__ orb_off,
                                                       const double* __restrict__ evals, double* __restrict__ Y) {
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * n) return;
  const int i = t / n, j = t - (t / n) * n;
  const int nocc = L.mol_nocc[m];
  const double* e = evals + orb_off[m];
  double* y = Y + L.mol_mat_off[m] + t;
  const bool oi = i < nocc, oj = j < nocc;
  *y = oi == oj ? 0.0 : *y * 2.0 / (oi ? e[i] - e[j] : e[j] - e[i]);
}

// The chemical-potential shift that keeps the trace: P1 += diag(dp/dmu) dmu1,
// dmu1 = -tr P1 / sum dp/dmu, dp/dmu = beta p0 (1 - p0) of the final p0; one block per molecule.
__global__ __launch_bounds__(256) void k_canon_shift(pyseqm_layout_t L, const int* __restrict__ idx,
                                                     const long long* __restrict__ orb_off,
                                                     const double* __restrict__ evals, const double* __restrict__ mu,
                                                     double beta, int steps, double* __restrict__ Y) {
  __shared__ double red[4];
  __shared__ double dp[256];
  const int m = idx[blockIdx.x], tid = threadIdx.x;
  const int n = L.mol_norb[m];
  double* Ym = Y + L.mol_mat_off[m];
  const double* e = evals + orb_off[m];
  const double cn = ldexp(beta, -2 - steps);
  double tr = 0.0, sdp = 0.0;
  for (int i = tid; i < n; i += 256) {
    double p0 = 0.5 - cn * (e[i] - mu[m]);
    for (int s = 0; s < steps; ++s) {
      const double p02 = p0 * p0;
      p0 = p02 / (2.0 * (p02 - p0) + 1.0) ;
    }
    dp[i] = beta * p0 * (1.0 - p0);
    tr += Ym[(size_t)i * n + i];
    sdp += dp[i];
  }
  tr = block_sum256(tr, red);
  sdp = block_sum256(sdp, red);
  const double dmu1 = -tr / sdp;
  for (int i = tid; i < n; i += 256) Ym[(size_t)i * n + i] += dp[i] * dmu1;
}

// out[slot] = sum over the block of A * B (Frobenius product), one block per molecule
__global__ __launch_bounds__(256) void k_block_dot(pyseqm_layout_t L, const int* __restrict__ idx,
                                                   const double* __restrict__ A, const double* __restrict__ B,
                                                   double* __restrict__ out) {
  __shared__ double red[4];
  const int m = idx[blockIdx.x];
  const int n = L.mol_norb[m];
  const size_t mo = L.mol_mat_off[m];
  double s = 0.0;
  for (int t = threadIdx.x; t < n * n; t += 256) s += A[mo + t] * B[mo + t];
  s = block_sum256(s, red);
  if (threadIdx.x == 0) out[m] = s;
}

// Y <- Y + s c_m X (mode 0), Y / sqrt(c_m) (1), s (Y - X) (2), s c_m X (3), Y + s X (4);
// c_m = c[m * cstride].
__global__ __launch_bounds__(256) void k_block_axpy(pyseqm_layout_t L, const int* __restrict__ idx, int mode,
                                                    double s, const double* __restrict__ c, int cstride,
                                                    const double* __restrict__ X, double* __restrict__ Y) {
  const int m = idx[blockIdx.y];
  const int n = L.mol_norb[m];
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * n) return;
  const size_t o = L.mol_mat_off[m] + t;
  const double cm = c ? c[(size_t)m * cstride] : 0.0;
  switch (mode) {
    case 0: Y[o] += s * cm * X[o]; break;
    case 1: Y[o] /= sqrt(cm); break;
    case 2: Y[o] = s * (Y[o] - X[o]); break;
    case 3: Y[o] = s * cm * X[o]; break;
    default: Y[o] += s * X[o]; break;
  }
}

// alpha = O^{-1} rhs per molecule (Gaussian elimination with partial pivoting, r <= kMaxRank);
// O[m][i][j] at m * kMaxRank^2 + i * kMaxRank + j, rhs / alpha at m * kMaxRank + i.
__global__ __launch_bounds__(64) void k_small_solve(const int* __restrict__ idx, int nact, int r,
                                                    const double* __restrict__ O, const double* __restrict__ rhs,
                                                    double* __restrict__ alpha) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= nact) return;
  const int m = idx[s];
  double a[kMaxRank][kMaxRank + 1];
  for (int i = 0; i < r; ++i) {
    for (int j = 0; j < r; ++j) a[i][j] = O[(size_t)m * kMaxRank * kMaxRank + i * kMaxRank + j];
    a[i][r] = rhs[(size_t)m * kMaxRank + i];
  }
  for (int k = 0; k < r; ++k) {
    int p = k;
    for (int i = k + 1; i < r; ++i)
      if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
    if (p != k)
      for (int j = k; j <= r; ++j) {
        const double t = a[k][j];
        a[k][j] = a[p][j];
        a[p][j] = t;
      }
    for (int i = k + 1; i < r; ++i) {
      const double l = a[i][k] / a[k][k];
      for (int j = k; j <= r; ++j) a[i][j] -= l * a[k][j];
    }
  }
  for (int i = r - 1; i >= 0; --i) {
    double v = a[i][r];
    for (int j = i + 1; j < r; ++j) v -= a[i][j] * alpha[(size_t)m * kMaxRank + j];
    alpha[(size_t)m * kMaxRank + i] = v / a[i][i];
  }
}

inline int elem_blocks(int nmax) { return (nmax * nmax + 255) / 256; }

}  // namespace

extern "C" int pyseqm_ksa_max_rank() { return kMaxRank; }

extern "C" long long pyseqm_fermi_workspace(int nact) {
  return nact < 0 ? -1 : (long long)nact * kFermiIter * (sizeof(double) + 1) + 16;
}

extern "C" int pyseqm_fermi_density(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                    const long long* orb_off, const double* evals, const double* Q, double T,
                                    double kB, double* occ, double* mu, double* D, double* S, void* work,
                                    void* stream) {
  if (!L || nact < 0 || nmax < 2 || nmax > 256 || !(T > 0.0) || !(kB > 0.0) || !work) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  const double beta = 1.0 / (kB * T);
  double* mu_hist = static_cast<double*>(work);
  unsigned char* conv = reinterpret_cast<unsigned char*>(mu_hist + (size_t)nact * kFermiIter);
  int* kpick = reinterpret_cast<int*>(conv + (((size_t)nact * kFermiIter + 7) & ~(size_t)7));
  k_fermi_newton<<<nact, 64, 0, s>>>(*L, idx, orb_off, evals, beta, mu_hist, conv);
  k_fermi_pick<<<1, 256, 0, s>>>(nact, conv, kpick);
  k_fermi_density<<<nact, 256, 0, s>>>(*L, idx, orb_off, evals, Q, beta, kB, mu_hist, kpick, occ, mu, D, S);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_block_gemm(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int ta, int tb,
                                 const double* A, const double* B, double* C, void* stream) {
  if (!L || nact < 0 || nmax < 1) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  const int nt = (nmax + 15) / 16;
  const dim3 grid(nt * nt, nact);
  hipStream_t s = (hipStream_t)stream;
  if (ta && tb)
    k_block_gemm<true, true><<<grid, 256, 0, s>>>(*L, idx, A, B, C);
  else if (ta)
    k_block_gemm<true, false><<<grid, 256, 0, s>>>(*L, idx, A, B, C);
  else if (tb)
    k_block_gemm<false, true><<<grid, 256, 0, s>>>(*L, idx, A, B, C);
  else
    k_block_gemm<false, false><<<grid, 256, 0, s>>>(*L, idx, A, B, C);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_canon_prt(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                const long long* orb_off, const double* evals, const double* mu, double T, double kB,
                                int steps, double* Y, void* stream) {
  if (!L || nact < 0 || nmax < 1 || nmax > 256 || steps < 0 || !(T > 0.0) || !(kB > 0.0)) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  const double beta = 1.0 / (kB * T);
  k_canon_elem<<<dim3(elem_blocks(nmax), nact), 256, 0, s>>>(*L, idx, orb_off, evals, mu, beta, steps, Y);
  k_canon_shift<<<nact, 256, 0, s>>>(*L, idx, orb_off, evals, mu, beta, steps, Y);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_density_response_elem(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                            const long long* orb_off, const double* evals, double* Y, void* stream) {
  if (!L || nact < 0 || nmax < 1) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  k_response_elem<<<dim3(elem_blocks(nmax), nact), 256, 0, (hipStream_t)stream>>>(*L, idx, orb_off, evals, Y);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_block_dot(const pyseqm_layout_t* L, const int* idx, int nact, const double* A,
                                const double* B, double* out, void* stream) {
  if (!L || nact < 0) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  k_block_dot<<<nact, 256, 0, (hipStream_t)stream>>>(*L, idx, A, B, out);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_block_axpy(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int mode, double s,
                                 const double* c, int cstride, const double* X, double* Y, void* stream) {
  if (!L || nact < 0 || nmax < 1 || mode < 0 || mode > 4 || ((mode == 0 || mode == 1 || mode == 3) && !c))
    return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  k_block_axpy<<<dim3(elem_blocks(nmax), nact), 256, 0, (hipStream_t)stream>>>(*L, idx, mode, s, c, cstride, X, Y);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_small_solve(const int* idx, int nact, int r, const double* O, const double* rhs, double* alpha,
                                  void* stream) {
  if (nact < 0 || r < 1 || r > kMaxRank) return PYSEQM_EINVAL;
  if (nact == 0) return PYSEQM_OK;
  k_small_solve<<<(nact + 63) / 64, 64, 0, (hipStream_t)stream>>>(idx, nact, r, O, rhs, alpha);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}
