// XL-BOMD and velocity-Verlet kernels on the packed batch, FP64, gfx950.
//
//   density propagation (MolecularDynamics.py:1418-1427, Niklasson et al. JCP 130, 214109, eq. 22):
//     P(n+1) = cD (c D(n) + (1 - c) P(n)) + sum_j coef[j] Pt[j],  Pt[slot] = P(n+1)
//     fused into one pass over the packed matrices (k + 3 reads, 2 writes per element);
//   XL electronic energy (energy.py:76-88): E = sum D F(P) - 1/2 sum (F(P) - H) P per molecule,
//     two-stage deterministic reduction;
//   velocity-Verlet half steps (MolecularDynamics.py:1449-1450, 1503-1505) on the packed atom
//     arrays (natoms, 3).
#include "common.h"

using namespace pyseqm;

namespace {

constexpr int kEl = 256;
constexpr int kMaxHist = 11;  // k <= 9 -> m = k + 1 <= 10 history matrices

struct XLCoef {
  double c[kMaxHist];
};

__global__ void k_xl_propagate(long long total, const double* __restrict__ D, double* __restrict__ P,
                               double* __restrict__ Pt, long long pt_stride, int m, XLCoef coef, double cD,
                               double c, int slot) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    double s = cD * (c * D[e] + (1.0 - c) * P[e]);
    for (int j = 0; j < m; ++j) s += coef.c[j] * Pt[(size_t)j * pt_stride + e];
    P[e] = s;
    Pt[(size_t)slot * pt_stride + e] = s;
  }
}

__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  __syncthreads();
  return s;
}

__global__ void k_xl_energy_partial(pyseqm_layout_t L, int nb, const double* __restrict__ D,
                                    const double* __restrict__ P, const double* __restrict__ F,
                                    const double* __restrict__ H, double* __restrict__ red) {
  __shared__ double sh[kEl / 64];
  const int m = blockIdx.y;
  const size_t n2 = (size_t)L.mol_norb[m] * L.mol_norb[m];
  const size_t off = L.mol_mat_off[m];
  double acc = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (size_t)nb * blockDim.x) {
    const double f = F[off + e];
    acc += D[off + e] * f - 0.5 * (f - H[off + e]) * P[off + e];
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) red[(size_t)m * nb + blockIdx.x] = s;
}

__global__ void k_xl_energy_final(int nmol, int nb, const double* __restrict__ red, double* __restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += red[(size_t)m * nb + b];
  out[m] = s;
}

// phase 0: v += dt/2 a; x += dt v.   phase 1: a = f * minv * acc_scale; v += dt/2 a.
__global__ void k_verlet(int natoms, int phase, double dt, double acc_scale, double* __restrict__ x,
                         double* __restrict__ v, double* __restrict__ a, const double* __restrict__ f,
                         const double* __restrict__ minv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * natoms) return;
  if (phase == 0) {
    const double vn = v[e] + 0.5 * dt * a[e];
    v[e] = vn;
    x[e] += dt * vn;
  } else {
    const double an = f[e] * minv[e / 3] * acc_scale;
    a[e] = an;
    v[e] += 0.5 * dt * an;
  }
}

inline int status() { return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH; }

}  // namespace

extern "C" int pyseqm_xl_propagate(long long total, const double* D, double* P, double* Pt, long long pt_stride,
                                   int m, const double* coef, double coef_D, double c, int slot, void* stream) {
  if (total < 0 || m < 1 || m > kMaxHist || slot < 0 || slot >= m || pt_stride < total || !coef) return PYSEQM_EINVAL;
  if (total == 0) return PYSEQM_OK;
  XLCoef cf{};
  for (int j = 0; j < m; ++j) cf.c[j] = coef[j];
  long long nb = (total + kEl - 1) / kEl;
  if (nb > 8192) nb = 8192;
  k_xl_propagate<<<(int)nb, kEl, 0, (hipStream_t)stream>>>(total, D, P, Pt, pt_stride, m, cf, coef_D, c, slot);
  return status();
}

extern "C" int pyseqm_xl_energy(const pyseqm_layout_t* L, const double* D, const double* P, const double* F,
                                const double* H, double* eelec, double* red, int nb, void* stream) {
  if (!L || nb < 1 || nb > PYSEQM_MAX_RED_BLOCKS) return PYSEQM_EINVAL;
  if (L->nmol == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_xl_energy_partial<<<dim3(nb, L->nmol), kEl, 0, s>>>(*L, nb, D, P, F, H, red);
  k_xl_energy_final<<<(L->nmol + 63) / 64, 64, 0, s>>>(L->nmol, nb, red, eelec);
  return status();
}

extern "C" int pyseqm_md_verlet(int natoms, int phase, double dt, double acc_scale, double* x, double* v, double* a,
                                const double* f, const double* minv, void* stream) {
  if (natoms < 0 || (phase != 0 && phase != 1) || (phase == 1 && (!f || !minv))) return PYSEQM_EINVAL;
  if (natoms == 0) return PYSEQM_OK;
  k_verlet<<<(3 * natoms + kEl - 1) / kEl, kEl, 0, (hipStream_t)stream>>>(natoms, phase, dt, acc_scale, x, v, a, f,
                                                                         minv);
  return status();
}
