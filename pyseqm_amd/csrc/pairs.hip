// Pair integrals for MNDO/AM1/PM3 sp atoms on gfx950.
//
// One 64-lane wave per 8x8 pair tile; each lane owns one atom pair and computes, in FP64:
//   * the local-frame two-electron integrals (MOPAC repp point-multipole formulas,
//     two_elec_two_center_int_local_frame.py:18-293),
//   * their rotation to the molecular frame, w = T L T^T with T the 10x10 map from local to
//     molecular orbital-product pairs (same rotation as two_elec_two_center_int.py:1384-1631),
//   * the electron-core attractions e1b/e2a (two_elec_two_center_int.py:1533-1569),
//   * the STO overlap block and the off-diagonal core Hamiltonian H_ij = S_ij (b_i + b_j)/2
//     (diat_overlap_PM6_SP.py:6-444, hcore.py:155-173),
//   * the core-core repulsion (energy.py:91-134).
// Diagonal core-Hamiltonian blocks are summed per atom from per-tile row/column partials in a
// fixed order (deterministic, no atomics).
#include "common.h"

using namespace pyseqm;

namespace {

__device__ constexpr double kTore[19] = {0, 1, 0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3, 4, 5, 6, 7, 0};
__device__ constexpr int kQn[19] = {0, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3};

// ---------------------------------------------------------------- per-atom setup
// rho1 = 1/(2 ad): hsp = ad/2 - 1/(2 sqrt(4 D1^2 + 1/ad^2)) solved by 5 secant steps from the
// closed-form start (cal_par.py:106-160).
__device__ double secant(int which, double target, double D, double x1) {
  double x2 = x1 + 0.04;
  for (int it = 0; it < 5; ++it) {
    double f1, f2;
    if (which == 1) {
      f1 = 0.5 * x1 - 0.5 / sqrt(4.0 * D * D + 1.0 / (x1 * x1));
      f2 = 0.5 * x2 - 0.5 / sqrt(4.0 * D * D + 1.0 / (x2 * x2));
    } else {
      f1 = 0.25 * x1 - 0.5 / sqrt(4.0 * D * D + 1.0 / (x1 * x1)) + 0.25 / sqrt(8.0 * D * D + 1.0 / (x1 * x1));
      f2 = 0.25 * x2 - 0.5 / sqrt(4.0 * D * D + 1.0 / (x2 * x2)) + 0.25 / sqrt(8.0 * D * D + 1.0 / (x2 * x2));
    }
    const double x3 = fabs(f2 - f1) > 1.0e-16 ? x1 + (x2 - x1) * (target - f1) / (f2 - f1) : x2;
    x1 = x2;
    x2 = x3;
  }
  return x2;
}

// The derived slots (dd, qq, rho0/1/2, tore, qn, Z) of an atom-parameter record from its inputs.
__device__ void derive_atom(double* p, int Z) {
  const double qn = kQn[Z];
  double dd = 0.0, qq = 0.0, r1 = 0.0, r2 = 0.0;
  if (Z > 2) {
    const double zs = p[PAR_ZS], zp = p[PAR_ZP];
    dd = (2.0 * qn + 1.0) * pow(4.0 * zs * zp, qn + 0.5) / pow(zs + zp, 2.0 * qn + 2.0) / sqrt(3.0);
    qq = sqrt((4.0 * qn * qn + 6.0 * qn + 2.0) / 20.0) / zp;
    const double hsp = p[PAR_HSP] / kEV;
    double d1 = cbrt(fabs(hsp) / (dd * dd));
    if (hsp < 0.0) d1 = -d1;
    r1 = 0.5 / secant(1, hsp, dd, d1);
    const double hpp = fmax(0.5 * (p[PAR_GPP] - p[PAR_GP2]), 0.1) / kEV;
    double q1 = pow(fabs(hpp) / 3.0 / (qq * qq * qq * qq), 0.2);
    if (hpp < 0.0) q1 = -q1;
    r2 = 0.5 / secant(2, hpp, qq, q1);
  }
  p[PAR_DD] = dd;
  p[PAR_QQ] = qq;
  p[PAR_RHO0] = 0.5 * kEV / p[PAR_GSS];
  p[PAR_RHO1] = r1;
  p[PAR_RHO2] = r2;
  p[PAR_TORE] = kTore[Z];
  p[PAR_QN] = qn;
  p[PAR_Z] = (double)Z;
}

__global__ void k_atom_setup(pyseqm_layout_t L, double* par) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= L.natoms) return;
  derive_atom(par + (size_t)a * PYSEQM_NPAR, L.atom_Z[a]);
}

// ------------------------------------------------------------ local-frame integrals
// 22 heavy-heavy integrals in MOPAC storage order (SS/SS)=0 ... (P*P/P*P)=21, local axis
// sigma pointing from atom j to atom i (two_elec_two_center_int_local_frame.py:126-262).
__device__ void local_xx(double r, double da, double db, double qa1, double qb1, double r0a,
                         double r0b, double r1a, double r1b, double r2a, double r2b, double ri[22]) {
  const double e1 = kEV * 0.5, e2 = kEV * 0.25, e3 = kEV * 0.125, e4 = kEV * 0.0625;
  const double qa = 2.0 * qa1, qb = 2.0 * qb1;
  const double aee = (r0a + r0b) * (r0a + r0b), ade = (r1a + r0b) * (r1a + r0b);
  const double aqe = (r2a + r0b) * (r2a + r0b), aed = (r0a + r1b) * (r0a + r1b);
  const double aeq = (r0a + r2b) * (r0a + r2b), axx = (r1a + r1b) * (r1a + r1b);
  const double adq = (r1a + r2b) * (r1a + r2b), aqd = (r2a + r1b) * (r2a + r1b);
  const double aqq = (r2a + r2b) * (r2a + r2b);
  const double r2 = r * r;
  auto inv = [](double x) { return 1.0 / sqrt(x); };
  const double ee = kEV * inv(r2 + aee);
  const double dze = -e1 * inv((r + da) * (r + da) + ade) + e1 * inv((r - da) * (r - da) + ade);
  const double c6 = e1 * inv(r2 + aqe);
  const double qzze = e2 * inv((r - qa) * (r - qa) + aqe) + e2 * inv((r + qa) * (r + qa) + aqe) - c6;
  const double qxxe = e1 * inv(r2 + qa * qa + aqe) - c6;
  const double edz = -e1 * inv((r - db) * (r - db) + aed) + e1 * inv((r + db) * (r + db) + aed);
  const double c12 = e1 * inv(r2 + aeq);
  const double eqzz = e2 * inv((r - qb) * (r - qb) + aeq) + e2 * inv((r + qb) * (r + qb) + aeq) - c12;
  const double eqxx = e1 * inv(r2 + qb * qb + aeq) - c12;
  const double t20 = e2 * inv((r + da) * (r + da) + adq), t22 = e2 * inv((r - da) * (r - da) + adq);
  const double t24 = e2 * inv((r - db) * (r - db) + aqd), t26 = e2 * inv((r + db) * (r + db) + aqd);
  const double t36 = e2 * inv(r2 + aqq), t39 = e2 * inv(r2 + qa * qa + aqq), t40 = e2 * inv(r2 + qb * qb + aqq);
  const double t42 = e3 * inv((r - qb) * (r - qb) + aqq), t44 = e3 * inv((r + qb) * (r + qb) + aqq);
  const double t46 = e3 * inv((r + qa) * (r + qa) + aqq), t48 = e3 * inv((r - qa) * (r - qa) + aqq);
  auto sq = [](double x) { return x * x; };
  ri[0] = ee;
  ri[1] = -dze;
  ri[2] = ee + qzze;
  ri[3] = ee + qxxe;
  ri[4] = -edz;
  ri[5] = e2 * inv(sq(r + da - db) + axx) + e2 * inv(sq(r - da + db) + axx) -
          e2 * inv(sq(r - da - db) + axx) - e2 * inv(sq(r + da + db) + axx);
  ri[6] = e1 * inv(r2 + sq(da - db) + axx) - e1 * inv(r2 + sq(da + db) + axx);
  ri[7] = -edz + e3 * inv(sq(r + qa - db) + aqd) - e3 * inv(sq(r + qa + db) + aqd) +
          e3 * inv(sq(r - qa - db) + aqd) - e3 * inv(sq(r - qa + db) + aqd) - t24 + t26;
  ri[8] = -edz - t24 + e2 * inv(sq(r - db) + qa * qa + aqd) + t26 - e2 * inv(sq(r + db) + qa * qa + aqd);
  ri[9] = e2 * inv(sq(qa1 - db) + sq(r + qa1) + aqd) - e2 * inv(sq(qa1 - db) + sq(r - qa1) + aqd) -
          e2 * inv(sq(qa1 + db) + sq(r + qa1) + aqd) + e2 * inv(sq(qa1 + db) + sq(r - qa1) + aqd);
  ri[10] = ee + eqzz;
  ri[11] = ee + eqxx;
  ri[12] = -dze + e3 * inv(sq(r + da - qb) + adq) - e3 * inv(sq(r - da - qb) + adq) +
           e3 * inv(sq(r + da + qb) + adq) - e3 * inv(sq(r - da + qb) + adq) + t22 - t20;
  ri[13] = -dze - t20 + e2 * inv(sq(r + da) + qb * qb + adq) + t22 - e2 * inv(sq(r - da) + qb * qb + adq);
  ri[14] = e2 * inv(sq(da - qb1) + sq(r - qb1) + adq) - e2 * inv(sq(da - qb1) + sq(r + qb1) + adq) -
           e2 * inv(sq(da + qb1) + sq(r - qb1) + adq) + e2 * inv(sq(da + qb1) + sq(r + qb1) + adq);
  ri[15] = ee + eqzz + qzze + e4 * inv(sq(r + qa - qb) + aqq) + e4 * inv(sq(r + qa + qb) + aqq) +
           e4 * inv(sq(r - qa - qb) + aqq) + e4 * inv(sq(r - qa + qb) + aqq) - t48 - t46 - t42 - t44 + t36;
  ri[16] = ee + eqzz + qxxe + e3 * inv(sq(r - qb) + qa * qa + aqq) + e3 * inv(sq(r + qb) + qa * qa + aqq) -
           t42 - t44 - t39 + t36;
  ri[17] = ee + eqxx + qzze + e3 * inv(sq(r + qa) + qb * qb + aqq) + e3 * inv(sq(r - qa) + qb * qb + aqq) -
           t46 - t48 - t40 + t36;
  const double qxxqxx = e3 * inv(r2 + sq(qa - qb) + aqq) + e3 * inv(r2 + sq(qa + qb) + aqq) - t39 - t40 + t36;
  ri[18] = ee + eqxx + qxxe + qxxqxx;
  const double dm = sq(qa1 - qb1), dp = sq(qa1 + qb1);
  ri[19] = e3 * inv(sq(r + qa1 - qb1) + dm + aqq) - e3 * inv(sq(r + qa1 + qb1) + dm + aqq) -
           e3 * inv(sq(r - qa1 - qb1) + dm + aqq) + e3 * inv(sq(r - qa1 + qb1) + dm + aqq) -
           e3 * inv(sq(r + qa1 - qb1) + dp + aqq) + e3 * inv(sq(r + qa1 + qb1) + dp + aqq) +
           e3 * inv(sq(r - qa1 - qb1) + dp + aqq) - e3 * inv(sq(r - qa1 + qb1) + dp + aqq);
  const double qxxqyy = e2 * inv(r2 + qa * qa + qb * qb + aqq) - t39 - t40 + t36;
  ri[20] = ee + eqxx + qxxe + qxxqyy;
  ri[21] = 0.5 * (qxxqxx - qxxqyy);
}

// 4 heavy-H integrals (two_elec_two_center_int_local_frame.py:90-107)
__device__ void local_xh(double r, double da, double qa1, double r0a, double r1a, double r2a,
                         double r0b, double ri[4]) {
  const double e1 = kEV * 0.5, e2 = kEV * 0.25;
  const double qa = 2.0 * qa1;
  const double aee = (r0a + r0b) * (r0a + r0b), ade = (r1a + r0b) * (r1a + r0b);
  const double aqe = (r2a + r0b) * (r2a + r0b);
  const double ee = kEV / sqrt(r * r + aee);
  const double c = e1 / sqrt(r * r + aqe);
  ri[0] = ee;
  ri[1] = e1 / sqrt((r + da) * (r + da) + ade) - e1 / sqrt((r - da) * (r - da) + ade);
  ri[2] = ee + e2 / sqrt((r + qa) * (r + qa) + aqe) + e2 / sqrt((r - qa) * (r - qa) + aqe) - c;
  ri[3] = ee + e1 / sqrt(r * r + qa * qa + aqe) - c;
}

// Rotation with R v = e_x built from the quaternion of v (two_elec_two_center_int.py:1576-1631),
// including its antipodal fallback (|1 + v_x| < 1e-7 -> constant 180 deg turn about z).
// mode 0: as the reference; 1: never take the fallback (smooth in v); 2: always the fallback.
// Finite differences use 1 or 2 chosen from the undisplaced geometry, which reproduces the
// reference's autograd derivative (zero frame derivative inside the fallback cone).
__device__ inline bool quat_flips(double vx) { return fabs(1.0 + vx) < 1.0e-7; }

__device__ inline void quat_frame(double vx, double vy, double vz, double R[3][3], int mode = 0) {
  double qy = vz, qz = -vy, qw = 1.0 + vx;
  if (mode == 2 || (mode == 0 && fabs(qw) < 1.0e-7)) {
    qy = 0.0;
    qz = 1.0;
    qw = 0.0;
  }
  const double n = 1.0 / sqrt(qy * qy + qz * qz + qw * qw);
  qy *= n;
  qz *= n;
  qw *= n;
  R[0][0] = 1.0 - 2.0 * (qy * qy + qz * qz);
  R[0][1] = -2.0 * (qz * qw);
  R[0][2] = 2.0 * (qy * qw);
  R[1][0] = 2.0 * (qz * qw);
  R[1][1] = 1.0 - 2.0 * (qz * qz);
  R[1][2] = 2.0 * (qy * qz);
  R[2][0] = -2.0 * (qy * qw);
  R[2][1] = 2.0 * (qy * qz);
  R[2][2] = 1.0 - 2.0 * (qy * qy);
}

// T[k][p]: coefficient of local pair p = (a>=b) in molecular pair k = (mu>=nu), where
// molecular p_m = sum_a R[a][m] * local_(1+a).
__device__ inline void pair_map(const double R[3][3], double T[10][10]) {
  double U[4][4];
  U[0][0] = 1.0;
#pragma unroll
  for (int a = 1; a < 4; ++a) U[0][a] = U[a][0] = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int a = 0; a < 3; ++a) U[1 + m][1 + a] = R[a][m];
#pragma unroll
  for (int k = 0; k < 10; ++k)
#pragma unroll
    for (int p = 0; p < 10; ++p) {
      const int mu = kTriI[k], nu = kTriJ[k], a = kTriI[p], b = kTriJ[p];
      T[k][p] = U[mu][a] * U[nu][b] + (a != b ? U[mu][b] * U[nu][a] : 0.0);
    }
}

// Local tensor over local orbital-product pairs (s,s)=0 (sig,s)=1 (sig,sig)=2 (pi,s)=3
// (pi,sig)=4 (pi,pi)=5 (pi',s)=6 (pi',sig)=7 (pi',pi)=8 (pi',pi')=9.
__device__ inline void local_tensor(const double ri[22], double Lm[10][10]) {
#pragma unroll
  for (int p = 0; p < 10; ++p)
#pragma unroll
    for (int q = 0; q < 10; ++q) Lm[p][q] = 0.0;
  Lm[0][0] = ri[0];  Lm[0][1] = ri[4];  Lm[0][2] = ri[10]; Lm[0][5] = ri[11]; Lm[0][9] = ri[11];
  Lm[1][0] = ri[1];  Lm[1][1] = ri[5];  Lm[1][2] = ri[12]; Lm[1][5] = ri[13]; Lm[1][9] = ri[13];
  Lm[2][0] = ri[2];  Lm[2][1] = ri[7];  Lm[2][2] = ri[15]; Lm[2][5] = ri[17]; Lm[2][9] = ri[17];
  Lm[3][3] = ri[6];  Lm[3][4] = ri[14];
  Lm[4][3] = ri[9];  Lm[4][4] = ri[19];
  Lm[5][0] = ri[3];  Lm[5][1] = ri[8];  Lm[5][2] = ri[16]; Lm[5][5] = ri[18]; Lm[5][9] = ri[20];
  Lm[6][6] = ri[6];  Lm[6][7] = ri[14];
  Lm[7][6] = ri[9];  Lm[7][7] = ri[19];
  Lm[8][8] = ri[21];
  Lm[9][0] = ri[3];  Lm[9][1] = ri[8];  Lm[9][2] = ri[16]; Lm[9][5] = ri[20]; Lm[9][9] = ri[18];
}

// ------------------------------------------------------------------- STO overlaps
// A_k(x) = int_1^inf t^k e^{-xt} dt and B_k(x) = int_-1^1 t^k e^{-xt} dt, k = 0..6, with the
// reference's recurrences / truncated series (diat_overlap_PM6_SP.py:464-670).
__device__ inline void aint7(double x, double a[7]) {
  if (x == 0.0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) a[k] = 0.0;
    return;
  }
  a[0] = exp(-x) / x;
#pragma unroll
  for (int k = 1; k < 7; ++k) a[k] = a[0] + k * a[k - 1] / x;
}

// Branch of bint7 at argument x: 2 recurrence (|x| > 0.5), 1 truncated series, 0 x ~ 0.
__device__ inline int bint_branch(double x) {
  const double ax = fabs(x);
  return ax > 0.5 ? 2 : (ax > 1.0e-6 ? 1 : 0);
}

// forced >= 0 evaluates that branch regardless of x (finite differences of a pair energy keep the
// branch of the geometry they differentiate, as the reference's autograd does).
__device__ inline void bint7(double x, double b[7], int forced = -1) {
  const int br = forced >= 0 ? forced : bint_branch(x);
  if (br == 2) {
    const double tx = exp(x) / x, tmx = -exp(-x) / x;
    b[0] = tx + tmx;
#pragma unroll
    for (int k = 1; k < 7; ++k) b[k] = ((k & 1) ? -tx : tx) + tmx + k * b[k - 1] / x;
  } else if (br == 1) {
    const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2;
    b[0] = 2.0 + x2 / 3.0 + x4 / 60.0 + x6 / 2520.0;
    b[2] = 2.0 / 3.0 + x2 / 5.0 + x4 / 84.0 + x6 / 3240.0;
    b[4] = 2.0 / 5.0 + x2 / 7.0 + x4 / 108.0 + x6 / 3960.0;
    b[6] = 2.0 / 7.0 + x2 / 9.0 + x4 / 132.0 + x6 / 4680.0;
    b[1] = -2.0 / 3.0 * x - x3 / 15.0 - x5 / 420.0;
    b[3] = -2.0 / 5.0 * x - x3 / 21.0 - x5 / 540.0;
    b[5] = -2.0 / 7.0 * x - x3 / 27.0 - x5 / 660.0;
  } else {
    b[0] = 2.0; b[1] = 0.0; b[2] = 2.0 / 3.0; b[3] = 0.0; b[4] = 2.0 / 5.0; b[5] = 0.0; b[6] = 2.0 / 7.0;
  }
}

__device__ inline int jcall_of(int qi, int qj) {
  if (qi == 1 && qj == 1) return 2;
  if (qi == 2 && qj == 1) return 3;
  if (qi == 2 && qj == 2) return 4;
  if (qi == 3 && qj == 1) return 431;
  if (qi == 3 && qj == 2) return 5;
  if (qi == 3 && qj == 3) return 6;
  return 0;
}

// Branches of the B integrals for the four exponent pairs (s_i s_j, p_i s_j, s_i p_j, p_i p_j),
// 2 bits each.
__device__ inline int overlap_branches(double r, double zas, double zap, double zbs, double zbp) {
  return bint_branch(0.5 * r * (zas - zbs)) | (bint_branch(0.5 * r * (zap - zbs)) << 2) |
         (bint_branch(0.5 * r * (zas - zbp)) << 4) | (bint_branch(0.5 * r * (zap - zbp)) << 6);
}

// S[0..4] = S111 (s|s), S211 (p_sig(i)|s(j)), S121, S221, S222 in the diatomic frame
// (diat_overlap_PM6_SP.py:84-345).  za/zb = (zeta_s, zeta_p) of atoms i / j.
// bsel >= 0: B-integral branches (overlap_branches) to use instead of those of r.
__device__ void overlap_local(int jc, double r, double zas, double zap, double zbs, double zbp,
                              bool need_p_i, bool need_p_j, double S[5], int bsel = -1) {
  S[0] = S[1] = S[2] = S[3] = S[4] = 0.0;
  double A[7], B[7];
  const double s3 = 1.7320508075688772;
  const double s10 = 3.1622776601683795, s30 = 5.477225575051661;
  auto AB = [&](double z1, double z2) {
    const int ci = (z1 == zas ? 0 : 1) | (z2 == zbs ? 0 : 2);
    aint7(0.5 * r * (z1 + z2), A);
    bint7(0.5 * r * (z1 - z2), B, bsel >= 0 ? (bsel >> (2 * ci)) & 3 : -1);
  };
  if (jc == 2) {
    AB(zas, zbs);
    S[0] = pow(zas * zbs * r * r, 1.5) * (A[2] * B[0] - B[2] * A[0]) / 4.0;
  } else if (jc == 3) {
    const double r4 = r * r * r * r;
    AB(zas, zbs);
    S[0] = pow(zbs, 1.5) * pow(zas, 2.5) * r4 * (A[3] * B[0] - B[3] * A[0] + A[2] * B[1] - B[2] * A[1]) / (s3 * 8.0);
    if (need_p_i) {
      AB(zap, zbs);
      S[1] = pow(zbs, 1.5) * pow(zap, 2.5) * r4 * (A[2] * B[0] - B[2] * A[0] + A[3] * B[1] - B[3] * A[1]) / 8.0;
    }
  } else if (jc == 4) {
    const double r5 = r * r * r * r * r;
    AB(zas, zbs);
    S[0] = pow(zbs * zas, 2.5) * r5 * (A[4] * B[0] + B[4] * A[0] - 2.0 * A[2] * B[2]) / 48.0;
    AB(zap, zbs);
    S[1] = pow(zbs * zap, 2.5) * r5 *
           (A[3] * (B[0] - B[2]) - A[1] * (B[2] - B[4]) + B[3] * (A[0] - A[2]) - B[1] * (A[2] - A[4])) /
           (16.0 * s3);
    AB(zas, zbp);
    S[2] = pow(zbp * zas, 2.5) * r5 *
           (A[3] * (B[0] - B[2]) - A[1] * (B[2] - B[4]) - B[3] * (A[0] - A[2]) + B[1] * (A[2] - A[4])) /
           (16.0 * s3);
    AB(zap, zbp);
    const double w = pow(zbp * zap, 2.5) * r5 / 16.0;
    S[3] = -w * (B[2] * (A[4] + A[0]) - A[2] * (B[4] + B[0]));
    S[4] = 0.5 * w * (A[4] * (B[0] - B[2]) - B[4] * (A[0] - A[2]) - A[2] * B[0] + B[2] * A[0]);
  } else if (jc == 431) {
    const double r5 = r * r * r * r * r;
    AB(zas, zbs);
    S[0] = pow(zbs, 1.5) * pow(zas, 3.5) * r5 *
           (A[4] * B[0] + 2.0 * B[1] * A[3] - 2.0 * A[1] * B[3] - B[4] * A[0]) / (s10 * 24.0);
    if (need_p_i) {
      AB(zap, zbs);
      S[1] = pow(zbs, 1.5) * pow(zap, 3.5) * r5 *
             (A[3] * (B[0] + B[2]) - A[1] * (B[4] + B[2]) + B[1] * (A[2] + A[4]) - B[3] * (A[2] + A[0])) /
             (8.0 * s30);
    }
  } else if (jc == 5) {
    const double r6 = r * r * r * r * r * r;
    AB(zas, zbs);
    S[0] = pow(zbs, 2.5) * pow(zas, 3.5) * r6 *
           (A[5] * B[0] + B[1] * A[4] - 2.0 * B[2] * A[3] - 2.0 * A[2] * B[3] + B[4] * A[1] + B[5] * A[0]) /
           (s30 * 48.0);
    AB(zap, zbs);
    S[1] = pow(zbs, 2.5) * pow(zap, 3.5) * r6 *
           (A[4] * B[0] + B[1] * A[5] - 2.0 * B[3] * A[3] - 2.0 * A[2] * B[2] + A[1] * B[5] + A[0] * B[4]) /
           (48.0 * s10);
    AB(zas, zbp);
    S[2] = pow(zbp, 2.5) * pow(zas, 3.5) * r6 *
           ((A[4] * B[0] - A[5] * B[1]) + 2.0 * (A[3] * B[1] - A[4] * B[2]) - 2.0 * (A[1] * B[3] - A[2] * B[4]) -
            (A[0] * B[4] - A[1] * B[5])) /
           (48.0 * s10);
    AB(zap, zbp);
    const double pre = pow(zbp, 2.5) * pow(zap, 3.5) * r6;
    S[3] = pre *
           ((A[3] * B[0] - A[5] * B[2]) + (A[2] * B[1] - A[4] * B[3]) - (A[1] * B[2] - A[3] * B[4]) -
            (A[0] * B[3] - A[2] * B[5])) /
           (16.0 * s30);
    S[4] = pre *
           ((A[5] - A[3]) * (B[0] - B[2]) + (A[4] - A[2]) * (B[1] - B[3]) - (A[3] - A[1]) * (B[2] - B[4]) -
            (A[2] - A[0]) * (B[3] - B[5])) /
           (32.0 * s30);
  } else if (jc == 6) {
    const double r7 = r * r * r * r * r * r * r;
    AB(zas, zbs);
    S[0] = pow(zbs * zas, 3.5) * r7 * (A[6] * B[0] - 3.0 * B[2] * A[4] + 3.0 * A[2] * B[4] - A[0] * B[6]) / 1440.0;
    AB(zap, zbs);
    S[1] = pow(zbs * zap, 3.5) * r7 *
           ((A[5] * B[0] + A[6] * B[1]) + (-A[4] * B[1] - A[5] * B[2]) - 2.0 * (A[3] * B[2] + A[4] * B[3]) -
            2.0 * (-A[2] * B[3] - A[3] * B[4]) + (A[1] * B[4] + A[2] * B[5]) + (-A[0] * B[5] - A[1] * B[6])) /
           (480.0 * s3);
    AB(zas, zbp);
    S[2] = pow(zbp * zas, 3.5) * r7 *
           ((A[5] * B[0] - A[6] * B[1]) + (A[4] * B[1] - A[5] * B[2]) - 2.0 * (A[3] * B[2] - A[4] * B[3]) -
            2.0 * (A[2] * B[3] - A[3] * B[4]) + (A[1] * B[4] - A[2] * B[5]) + (A[0] * B[5] - A[1] * B[6])) /
           (480.0 * s3);
    AB(zap, zbp);
    const double pre = pow(zbp, 3.5) * pow(zap, 3.5) * r7;
    S[3] = pre * ((A[4] * B[0] - A[6] * B[2]) - 2.0 * (A[2] * B[2] - A[4] * B[4]) + (A[0] * B[4] - A[2] * B[6])) / 480.0;
    S[4] = pre *
           ((A[6] - A[4]) * (B[0] - B[2]) - 2.0 * (A[4] - A[2]) * (B[2] - B[4]) + (A[2] - A[0]) * (B[4] - B[6])) /
           960.0;
  }
  (void)need_p_j;
}

// --------------------------------------------------------------- core-core repulsion
__device__ double core_repulsion(int method, int Zi, int Zj, double rbohr, double gam, const double* pi,
                            const double* pj) {
  const double ra = rbohr * kA0;
  const double ti = pi[PAR_TORE], tj = pj[PAR_TORE];
  const double t1 = ti * tj * gam;
  double t2 = exp(-pi[PAR_ALPHA] * ra);
  if ((Zi == 7 || Zi == 8) && Zj == 1) t2 *= ra;
  const double t3 = exp(-pj[PAR_ALPHA] * ra);
  double e = t1 * (1.0 + t2 + t3);
  const int ng = method == PYSEQM_AM1 ? 4 : (method == PYSEQM_PM3 ? 2 : 0);
  if (ng) {
    double g = 0.0;
    for (int k = 0; k < ng; ++k) {
      const double* ki = pi + PAR_GAUSS + 3 * k;
      const double* kj = pj + PAR_GAUSS + 3 * k;
      g += ki[0] * exp(-ki[1] * (ra - ki[2]) * (ra - ki[2]));
      g += kj[0] * exp(-kj[1] * (ra - kj[2]) * (ra - kj[2]));
    }
    e += ti * tj / ra * g;
  }
  return e;
}

struct PairGeom {
  double x[3];  // unit vector i -> j
  double r;     // bohr
};

__device__ inline PairGeom geom_from_vector(double dx, double dy, double dz) {
  PairGeom g;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  g.x[0] = dx / d;
  g.x[1] = dy / d;
  g.x[2] = dz / d;
  g.r = d * (1.0 / kA0);
  return g;
}

__device__ inline PairGeom pair_geom(const double* coords, int i, int j) {
  return geom_from_vector(coords[3 * j] - coords[3 * i], coords[3 * j + 1] - coords[3 * i + 1],
                          coords[3 * j + 2] - coords[3 * i + 2]);
}

// Molecular-frame two-electron integrals of a heavy-heavy pair, column by column:
// fn(k, l, (mu nu|la si)) for k = tril(mu, nu) on atom i and l = tril(la, si) on atom j.
template <class Fn>
__device__ inline void xx_integrals(const PairGeom& g, const double* pi, const double* pj, Fn&& fn, int fmode = 0) {
  double ri[22];
  local_xx(g.r, pi[PAR_DD], pj[PAR_DD], pi[PAR_QQ], pj[PAR_QQ], pi[PAR_RHO0], pj[PAR_RHO0], pi[PAR_RHO1],
           pj[PAR_RHO1], pi[PAR_RHO2], pj[PAR_RHO2], ri);
  double R[3][3];
  quat_frame(-g.x[0], -g.x[1], -g.x[2], R, fmode);
  double T[10][10], Lm[10][10];
  pair_map(R, T);
  local_tensor(ri, Lm);
#pragma unroll
  for (int l = 0; l < 10; ++l) {
    double v[10];
#pragma unroll
    for (int p = 0; p < 10; ++p) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 10; ++q) s += Lm[p][q] * T[l][q];
      v[p] = s;
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      double s = 0.0;
#pragma unroll
      for (int p = 0; p < 10; ++p) s += T[k][p] * v[p];
      fn(k, l, s);
    }
  }
}

// Heavy-H integrals (mu nu|s s), k = tril(mu, nu) on the heavy atom i.
template <class Fn>
__device__ inline void xh_integrals(const PairGeom& g, const double* pi, const double* pj, Fn&& fn, int fmode = 0) {
  double ri[4];
  local_xh(g.r, pi[PAR_DD], pi[PAR_QQ], pi[PAR_RHO0], pi[PAR_RHO1], pi[PAR_RHO2], pj[PAR_RHO0], ri);
  double R[3][3];
  quat_frame(-g.x[0], -g.x[1], -g.x[2], R, fmode);
  double T[10][10];
  pair_map(R, T);
#pragma unroll
  for (int k = 0; k < 10; ++k) fn(k, T[k][0] * ri[0] + T[k][1] * ri[1] + T[k][2] * ri[2] + (T[k][5] + T[k][9]) * ri[3]);
}

__device__ inline double hh_integral(const PairGeom& g, const double* pi, const double* pj) {
  return kEV / sqrt(g.r * g.r + (pi[PAR_RHO0] + pj[PAR_RHO0]) * (pi[PAR_RHO0] + pj[PAR_RHO0]));
}

// Off-diagonal core block H_ij[a][b] = S_ij[a][b] (beta_a(i) + beta_b(j)) / 2 for NI x NJ orbitals.
template <int NI, int NJ>
__device__ inline void core_block(const PairGeom& g, const double* pi, const double* pj, double Hb[4][4],
                                  int fmode = 0, int bsel = -1) {
  double di[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) di[a][b] = 0.0;
  if (g.r <= kOverlapCut) {
    const int jc = jcall_of((int)pi[PAR_QN], (int)pj[PAR_QN]);  // validated on the host
    double S[5];
    overlap_local(jc, g.r, pi[PAR_ZS], pi[PAR_ZP], pj[PAR_ZS], pj[PAR_ZP], NI == 4, NJ == 4, S, bsel);
    double Rp[3][3];
    quat_frame(g.x[0], g.x[1], g.x[2], Rp, fmode);
    di[0][0] = S[0];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      di[1 + m][0] = S[1] * Rp[0][m];
      di[0][1 + m] = -S[2] * Rp[0][m];
    }
    if (NJ == 4) {
      const double D[3] = {-S[3], S[4], S[4]};
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n)
          di[1 + m][1 + n] = Rp[0][m] * D[0] * Rp[0][n] + Rp[1][m] * D[1] * Rp[1][n] + Rp[2][m] * D[2] * Rp[2][n];
    }
  }
  const double bi[4] = {pi[PAR_BS], pi[PAR_BP], pi[PAR_BP], pi[PAR_BP]};
  const double bj[4] = {pj[PAR_BS], pj[PAR_BP], pj[PAR_BP], pj[PAR_BP]};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Hb[a][b] = (a < NI && b < NJ) ? di[a][b] * 0.5 * (bi[a] + bj[b]) : 0.0;
}

// ------------------------------------------------------------------- tile kernels
enum { XX = 0, XH = 1, HH = 2 };

template <int KIND>
__global__ __launch_bounds__(64) void k_pair_tiles(pyseqm_layout_t L, int method, int tile_base,
                                                    const double* __restrict__ coords,
                                                    const double* __restrict__ par, double* __restrict__ wout,
                                                    double* __restrict__ H, double* __restrict__ part,
                                                    double* __restrict__ enuc_tile,
                                                    double* __restrict__ enuc_pair,
                                                    const int* __restrict__ pair_ref) {
  const int tile = tile_base + xcd_run(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const Tile t = load_tile(L.tiles, tile);
  int i, j;
  const bool valid = lane_pair(t, lane, i, j);
  constexpr int NR = KIND == HH ? 1 : 10;
  constexpr int NC = KIND == XX ? 10 : 1;
  constexpr int NI = KIND == HH ? 1 : 4;
  constexpr int NJ = KIND == XX ? 4 : 1;
  double rowv[NR], colv[NC];
#pragma unroll
  for (int k = 0; k < NR; ++k) rowv[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NC; ++k) colv[k] = 0.0;
  double en = 0.0;
  constexpr int NW = KIND == XX ? 100 : (KIND == XH ? 10 : 1);
  const int S = tile_wstride(t);
  double* wt = wout + L.tile_woff[tile] + (valid ? lane_rank(t, lane) : 0);

  if (valid) {
    const double* pi = par + (size_t)i * PYSEQM_NPAR;
    const double* pj = par + (size_t)j * PYSEQM_NPAR;
    const PairGeom g = pair_geom(coords, i, j);
    const double ti = pi[PAR_TORE], tj = pj[PAR_TORE];
    double gam = 0.0;
    if constexpr (KIND == XX) {
      xx_integrals(g, pi, pj, [&](int k, int l, double w) {
        wt[(size_t)(k * 10 + l) * S] = w;
        if (l == 0) rowv[k] = -tj * w;  // e1b: electrons of i, core of j
        if (k == 0) colv[l] = -ti * w;  // e2a: electrons of j, core of i
        if (k == 0 && l == 0) gam = w;
      });
    } else if constexpr (KIND == XH) {
      xh_integrals(g, pi, pj, [&](int k, double w) {
        wt[(size_t)k * S] = w;
        rowv[k] = -tj * w;
        if (k == 0) gam = w;
      });
      colv[0] = -ti * gam;
    } else {
      gam = hh_integral(g, pi, pj);
      wt[0] = gam;
      rowv[0] = -tj * gam;
      colv[0] = -ti * gam;
    }
    double Hb[4][4];
    core_block<NI, NJ>(g, pi, pj, Hb);
    const long long mo = L.mol_mat_off[t.mol];
    const int norb = L.mol_norb[t.mol];
    const int ai = L.atom_ao[i], aj = L.atom_ao[j];
    double* Hm = H + mo;
#pragma unroll
    for (int a = 0; a < NI; ++a)
#pragma unroll
      for (int b = 0; b < NJ; ++b) {
        Hm[(size_t)(ai + a) * norb + aj + b] = Hb[a][b];
        Hm[(size_t)(aj + b) * norb + ai + a] = Hb[a][b];
      }
    en = core_repulsion(method, (int)pi[PAR_Z], (int)pj[PAR_Z], g.r, gam, pi, pj);
    if (enuc_pair) enuc_pair[pair_ref[(size_t)tile * 64 + lane]] = en;
  }
  write_partials_c(part, L, tile, lane, rowv, colv);
  const double es = wave64(en);  // every lane active here
  if (lane == 0) enuc_tile[tile] = es;
}

// Hellmann-Feynman pair energy at fixed density: the part of 0.5 sum P (H + F) + E_nuc that
// depends on the geometry of pair (i, j) (basics.py:1315-1336 differentiates the same energy with
// scf_backward = 0): 2 sum P_ij H_ij + sum P_ii e1b + sum P_jj e2a + sum P_ii P_jj (..|..)
// - 1/2 sum P_ij P_ij (..|..) + E_nuc(i, j).
// With XL = true the energy is the extended-Lagrangian one, sum D F(P) - 1/2 (F(P) - H) P + E_nuc
// (energy.py:76-88, differentiated at fixed D and P by xlbomd.py:536-551): H and the electron-core
// terms contract with D, the two-electron terms are B(D, P) - 1/2 B(P, P) for the symmetric
// bilinear form B(X, Y) = sum X G(Y); D = P gives the energy above.
struct PairDensity {
  double pa[10], pb[10];  // diagonal blocks, tril, off-diagonal entries doubled
  double pab[4][4];       // block (i, j)
};

// MODE 0: the pair's share of Eelec(P) + E_nuc; 1 (XL-BOMD): of sum D F(P) - 1/2 sum (F(P) - H) P +
// E_nuc; 2 (linear): of sum D F(P), no core repulsion (rep_out = 0).
template <int KIND, int MODE>
__device__ double pair_energy(int method, const PairGeom& g, const double* pi, const double* pj,
                              const PairDensity& P, const PairDensity& D, int wmode, int smode, int bsel,
                              double* rep_out = nullptr) {
  constexpr int NI = KIND == HH ? 1 : 4;
  constexpr int NJ = KIND == XX ? 4 : 1;
  const double ti = pi[PAR_TORE], tj = pj[PAR_TORE];
  double e = 0.0, gam = 0.0;
  // Coulomb weight of w_kl: P_a P_b (plain), D_a P_b + P_a D_b - P_a P_b (XL), D_a P_b + P_a D_b (linear)
  auto coul = [&](int k, int l) {
    return MODE == 1   ? D.pa[k] * P.pb[l] + P.pa[k] * D.pb[l] - P.pa[k] * P.pb[l]
           : MODE == 2 ? D.pa[k] * P.pb[l] + P.pa[k] * D.pb[l]
                       : P.pa[k] * P.pb[l];
  };
  // exchange weight: 1/2 P_ab (plain), D_ab - 1/2 P_ab (XL), D_ab (linear), times K(P)_ab
  auto xw = [&](int a, int b) {
    return MODE == 1 ? D.pab[a][b] - 0.5 * P.pab[a][b] : MODE == 2 ? D.pab[a][b] : 0.5 * P.pab[a][b];
  };
  if constexpr (KIND == XX) {
    double K[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) K[a][b] = 0.0;
    xx_integrals(g, pi, pj, [&](int k, int l, double w) {
      const int mu = kTriI[k], nu = kTriJ[k], la = kTriI[l], si = kTriJ[l];
      e += coul(k, l) * w;
      if (l == 0) e -= tj * D.pa[k] * w;
      if (k == 0) e -= ti * D.pb[l] * w;
      if (k == 0 && l == 0) gam = w;
      K[mu][la] += w * P.pab[nu][si];
      if (mu != nu) K[nu][la] += w * P.pab[mu][si];
      if (la != si) K[mu][si] += w * P.pab[nu][la];
      if (mu != nu && la != si) K[nu][si] += w * P.pab[mu][la];
    }, wmode);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) e -= xw(a, b) * K[a][b];
  } else if constexpr (KIND == XH) {
    double K[4] = {0.0, 0.0, 0.0, 0.0};
    xh_integrals(g, pi, pj, [&](int k, double w) {
      const int mu = kTriI[k], nu = kTriJ[k];
      e += coul(k, 0) * w - tj * D.pa[k] * w;
      if (k == 0) {
        gam = w;
        e -= ti * D.pb[0] * w;
      }
      K[mu] += w * P.pab[nu][0];
      if (mu != nu) K[nu] += w * P.pab[mu][0];
    }, wmode);
#pragma unroll
    for (int a = 0; a < 4; ++a) e -= xw(a, 0) * K[a];
  } else {
    gam = hh_integral(g, pi, pj);
    e = (coul(0, 0) - tj * D.pa[0] - ti * D.pb[0] - xw(0, 0) * P.pab[0][0]) * gam;
  }
  double Hb[4][4];
  core_block<NI, NJ>(g, pi, pj, Hb, smode, bsel);
#pragma unroll
  for (int a = 0; a < NI; ++a)
#pragma unroll
    for (int b = 0; b < NJ; ++b) e += 2.0 * D.pab[a][b] * Hb[a][b];
  const double rep = MODE == 2 ? 0.0 : core_repulsion(method, (int)pi[PAR_Z], (int)pj[PAR_Z], g.r, gam, pi, pj);
  if (rep_out) *rep_out = rep;
  return e + rep;
}

template <int NI, int NJ>
__device__ void load_pair_density(const double* Pm, int n, int ai, int aj, PairDensity& D) {
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const int mu = kTriI[k], nu = kTriJ[k];
    const double f = mu == nu ? 1.0 : 2.0;
    D.pa[k] = (mu < NI && nu < NI) ? f * Pm[(size_t)(ai + mu) * n + ai + nu] : 0.0;
    D.pb[k] = (mu < NJ && nu < NJ) ? f * Pm[(size_t)(aj + mu) * n + aj + nu] : 0.0;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) D.pab[a][b] = (a < NI && b < NJ) ? Pm[(size_t)(ai + a) * n + aj + b] : 0.0;
}

// Forces: dE_pair/dR (R = r_j - r_i, Angstrom) by a 4th-order central difference of the pair
// energy at fixed density (step h); +dE/dR on atom i, -dE/dR on atom j, reduced per tile.
// XL: Dm is the density contracted with H (the XL-BOMD D), P the propagated field.
template <int KIND, bool XL>
__global__ __launch_bounds__(64) void k_pair_forces(pyseqm_layout_t L, int method, int tile_base,
                                                     const double* __restrict__ coords,
                                                     const double* __restrict__ par, const double* __restrict__ P,
                                                     const double* __restrict__ Dm, double h,
                                                     double* __restrict__ part) {
  const int tile = tile_base + xcd_run(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const Tile t = load_tile(L.tiles, tile);
  int i, j;
  const bool valid = lane_pair(t, lane, i, j);
  double rowv[3] = {0.0, 0.0, 0.0}, colv[3] = {0.0, 0.0, 0.0};
  if (valid) {
    const double* pi = par + (size_t)i * PYSEQM_NPAR;
    const double* pj = par + (size_t)j * PYSEQM_NPAR;
    const long long mo = L.mol_mat_off[t.mol];
    const int n = L.mol_norb[t.mol];
    const int ai = L.atom_ao[i], aj = L.atom_ao[j];
    constexpr int NI = KIND == HH ? 1 : 4;
    constexpr int NJ = KIND == XX ? 4 : 1;
    PairDensity Pd, Dd;
    load_pair_density<NI, NJ>(P + mo, n, ai, aj, Pd);
    if constexpr (XL) load_pair_density<NI, NJ>(Dm + mo, n, ai, aj, Dd);
    const PairDensity& Dr = XL ? Dd : Pd;
    const double R0[3] = {coords[3 * j] - coords[3 * i], coords[3 * j + 1] - coords[3 * i + 1],
                          coords[3 * j + 2] - coords[3 * i + 2]};
    const double c1 = 8.0 / (12.0 * h), c2 = -1.0 / (12.0 * h);
    const PairGeom g0 = geom_from_vector(R0[0], R0[1], R0[2]);
    const int wmode = quat_flips(-g0.x[0]) ? 2 : 1;  // frame of -x_ij (integrals)
    const int smode = quat_flips(g0.x[0]) ? 2 : 1;   // frame of +x_ij (overlaps)
    const int bsel = overlap_branches(g0.r, pi[PAR_ZS], pi[PAR_ZP], pj[PAR_ZS], pj[PAR_ZP]);
#pragma unroll 1
    for (int x = 0; x < 3; ++x) {
      double gx = 0.0;
#pragma unroll 1
      for (int sgn = -2; sgn <= 2; ++sgn) {
        if (sgn == 0) continue;
        double R[3] = {R0[0], R0[1], R0[2]};
        R[x] += sgn * h;
        const double e =
            pair_energy<KIND, XL>(method, geom_from_vector(R[0], R[1], R[2]), pi, pj, Pd, Dr, wmode, smode, bsel);
        const double c = (sgn == 1 || sgn == -1) ? c1 : c2;
        gx += (sgn > 0 ? c : -c) * e;
      }
      rowv[x] = gx;    // force on i = -dE/dr_i = +dE/dR
      colv[x] = -gx;   // force on j
    }
  }
  write_partials(part, tile, lane, rowv, colv);
}

template <int G>
__global__ __launch_bounds__(256) void k_atom_forces(pyseqm_layout_t L, const double* __restrict__ part,
                                                     double* __restrict__ F) {
  const int a = blockIdx.x * (256 / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (a >= L.natoms) return;  // uniform over the lane group
  double f[3];
  gather_slots<G>(part, L.csr_idx, L.csr_off[a], L.csr_off[a + 1], lane, f);
  if (lane < 3) F[3 * a + lane] = lane == 0 ? f[0] : (lane == 1 ? f[1] : f[2]);
}

// ------------------------------------------------- energy gradient w.r.t. atom parameters
// At a converged closed-shell SCF density the energy is stationary in P, so dE/dtheta is the
// derivative at fixed P: the gradient the reference's autograd returns through H, the integrals and
// the core repulsion when the SCF density is detached (scf_backward = 0, scf_loop.py:1996-2031;
// learned parameters re-packed per call, basics.py:781-789).  Each pair's energy (its Eelec share and
// its core repulsion) depends on the input slots below of its two atoms, through the multipoles
// (dd, qq, rho0/1/2), the overlaps, the resonance integrals and the core-repulsion terms; a
// 4th-order central difference of the pair energy in each slot, with the overlap branches pinned at
// the unperturbed values, gives that pair's share.  U_ss, U_pp and g_sp reach the one-center terms
// only, which k_param_onecenter differentiates in closed form.
__device__ constexpr int kPairParams[21] = {PAR_ZS,        PAR_ZP,        PAR_BS,        PAR_BP,        PAR_GSS,
                                            PAR_GPP,       PAR_GP2,       PAR_HSP,       PAR_ALPHA,     PAR_GAUSS,
                                            PAR_GAUSS + 1, PAR_GAUSS + 2, PAR_GAUSS + 3, PAR_GAUSS + 4, PAR_GAUSS + 5,
                                            PAR_GAUSS + 6, PAR_GAUSS + 7, PAR_GAUSS + 8, PAR_GAUSS + 9,
                                            PAR_GAUSS + 10, PAR_GAUSS + 11};
constexpr int kParamGroup = 5;  // slots per pass: 2 values (Eelec, core repulsion) each in a 10-wide partial
__host__ __device__ inline int pair_param_count(int method) {
  return method == PYSEQM_AM1 ? 21 : (method == PYSEQM_PM3 ? 15 : 9);
}
// Step of slot k of a record: hrel * max(|value|, 0.1).
__device__ inline double param_step(const double* p, int k, double hrel) { return hrel * fmax(fabs(p[k]), 0.1); }
__device__ constexpr int kStencil[4] = {-2, -1, 1, 2};
__device__ constexpr double kStencilW[4] = {1.0 / 12.0, -8.0 / 12.0, 8.0 / 12.0, -1.0 / 12.0};

// rec[((a * kParamGroup + g) * 4 + s) * NPAR ...]: atom a's record with slot kPairParams[p0 + g]
// moved by kStencil[s] steps and the derived slots recomputed.  One thread per (atom, g, s).
__global__ void k_param_records(pyseqm_layout_t L, const double* __restrict__ par, int p0, int np, double hrel,
                                double* __restrict__ rec) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = t & 3, g = (t >> 2) % kParamGroup, a = (t >> 2) / kParamGroup;
  if (a >= L.natoms || g >= np) return;
  const double* src = par + (size_t)a * PYSEQM_NPAR;
  double* q = rec + ((size_t)(a * kParamGroup + g) * 4 + s) * PYSEQM_NPAR;
  for (int k = 0; k < PYSEQM_NPAR; ++k) q[k] = src[k];
  const int slot = kPairParams[p0 + g];
  q[slot] = src[slot] + kStencil[s] * param_step(src, slot, hrel);
  derive_atom(q, L.atom_Z[a]);
}

// Per pair and slot of the pass: d(pair Eelec share)/d(slot) and d(core repulsion)/d(slot) for
// atom i (row partial values 2g, 2g + 1) and for atom j (column partial); MODE 2: the pair's share
// of d(sum Y F(P))/d(slot) (no core repulsion) for a symmetric Y.
template <int KIND, int MODE>
__global__ __launch_bounds__(64) void k_pair_param_grad(pyseqm_layout_t L, int method, int tile_base,
                                                        const double* __restrict__ coords,
                                                        const double* __restrict__ par, const double* __restrict__ P,
                                                        const double* __restrict__ Y,
                                                        const double* __restrict__ rec, int p0, int np, double hrel,
                                                        double* __restrict__ part) {
  const int tile = tile_base + xcd_run(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const Tile t = load_tile(L.tiles, tile);
  int i, j;
  const bool valid = lane_pair(t, lane, i, j);
  double rowv[10], colv[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) rowv[k] = colv[k] = 0.0;
  if (valid) {
    const double* pi = par + (size_t)i * PYSEQM_NPAR;
    const double* pj = par + (size_t)j * PYSEQM_NPAR;
    const long long mo = L.mol_mat_off[t.mol];
    const int n = L.mol_norb[t.mol];
    constexpr int NI = KIND == HH ? 1 : 4;
    constexpr int NJ = KIND == XX ? 4 : 1;
    PairDensity Pd, Yd;
    load_pair_density<NI, NJ>(P + mo, n, L.atom_ao[i], L.atom_ao[j], Pd);
    if constexpr (MODE == 2) load_pair_density<NI, NJ>(Y + mo, n, L.atom_ao[i], L.atom_ao[j], Yd);
    const PairDensity& Dr = MODE == 2 ? Yd : Pd;
    const PairGeom g0 = pair_geom(coords, i, j);
    const int wmode = quat_flips(-g0.x[0]) ? 2 : 1;
    const int smode = quat_flips(g0.x[0]) ? 2 : 1;
    const int bsel = overlap_branches(g0.r, pi[PAR_ZS], pi[PAR_ZP], pj[PAR_ZS], pj[PAR_ZP]);
#pragma unroll 1
    for (int g = 0; g < np; ++g) {
      const int slot = kPairParams[p0 + g];
#pragma unroll 1
      for (int side = 0; side < 2; ++side) {
        const int a = side == 0 ? i : j;
        const double inv = 1.0 / param_step(side == 0 ? pi : pj, slot, hrel);
        double de = 0.0, dn = 0.0;
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
          const double* q = rec + ((size_t)(a * kParamGroup + g) * 4 + s) * PYSEQM_NPAR;
          double rep = 0.0;
          const double e = pair_energy<KIND, MODE>(method, g0, side == 0 ? q : pi, side == 0 ? pj : q, Pd, Dr,
                                                   wmode, smode, bsel, &rep);
          de += kStencilW[s] * (e - rep);
          dn += kStencilW[s] * rep;
        }
        if (side == 0) {
          rowv[2 * g] = de * inv;
          rowv[2 * g + 1] = dn * inv;
        } else {
          colv[2 * g] = de * inv;
          colv[2 * g + 1] = dn * inv;
        }
      }
    }
  }
  write_partials(part, tile, lane, rowv, colv);
}

// out[a][0][k] += sum of atom a's pair shares of dEelec/d(slot k), out[a][1][k] += the core
// repulsion's, for the pass's slots (fixed gather order).
template <int G>
__global__ __launch_bounds__(256) void k_param_gather(pyseqm_layout_t L, const double* __restrict__ part, int p0,
                                                      int np, double* __restrict__ out) {
  const int a = blockIdx.x * (256 / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (a >= L.natoms) return;  // uniform over the lane group
  double acc[10];
  gather_slots<G>(part, L.csr_idx, L.csr_off[a], L.csr_off[a + 1], lane, acc);
  if (lane < np) {
    double de = acc[0], dn = acc[1];
#pragma unroll
    for (int g = 1; g < kParamGroup; ++g)
      if (lane == g) {
        de = acc[2 * g];
        dn = acc[2 * g + 1];
      }
    const int slot = kPairParams[p0 + lane];
    out[(size_t)a * 48 + slot] += de;
    out[(size_t)a * 48 + 24 + slot] += dn;
  }
}

// The closed-form one-center part: E_A = sum_mu P_mumu U_mu + 1/2 sum P G1(P) over the atom's
// block (the one-center terms of the Fock build, fock.py:196-224), differentiated in U_ss, U_pp,
// g_ss, g_sp, g_pp, g_p2 and h_sp; with Y given, the same for sum Y (U + G1(P)) over the block.
// Writes every slot of out[a] (the pair passes add to it).
__global__ void k_param_onecenter(pyseqm_layout_t L, const double* __restrict__ P, const double* __restrict__ Y,
                                  double* __restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= L.natoms) return;
  const int m = L.atom_mol[a];
  const int n = L.mol_norb[m];
  const double* Pm = P + L.mol_mat_off[m];
  const int ao = L.atom_ao[a];
  double* o = out + (size_t)a * 48;
  for (int k = 0; k < 48; ++k) o[k] = 0.0;
  const double pss = Pm[(size_t)ao * n + ao];
  if (!Y) {
    o[PAR_USS] = pss;
    o[PAR_GSS] = 0.25 * pss * pss;
    if (L.atom_Z[a] > 1) {
      double pd[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) pd[k] = Pm[(size_t)(ao + kTriI[k]) * n + ao + kTriJ[k]];
      const double ppt = pd[2] + pd[5] + pd[9];
      const double psp2 = pd[1] * pd[1] + pd[3] * pd[3] + pd[6] * pd[6];
      const double ppd2 = pd[2] * pd[2] + pd[5] * pd[5] + pd[9] * pd[9];
      const double ppo2 = pd[4] * pd[4] + pd[7] * pd[7] + pd[8] * pd[8];
      o[PAR_UPP] = ppt;
      o[PAR_GSP] = pss * ppt - 0.5 * psp2;
      o[PAR_HSP] = -0.5 * pss * ppt + 1.5 * psp2;
      o[PAR_GPP] = 0.375 * ppd2 - 0.125 * ppt * ppt + 0.75 * ppo2;
      o[PAR_GP2] = 0.625 * (ppt * ppt - ppd2) - 1.25 * ppo2;
    }
    return;
  }
  // sum Y (U + G1(P)): diagonal elements once, off-diagonal pairs twice (both triangles)
  const double* Ym = Y + L.mol_mat_off[m];
  const double yss = Ym[(size_t)ao * n + ao];
  o[PAR_USS] = yss;
  o[PAR_GSS] = 0.5 * yss * pss;
  if (L.atom_Z[a] > 1) {
    double pd[10], yd[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      pd[k] = Pm[(size_t)(ao + kTriI[k]) * n + ao + kTriJ[k]];
      yd[k] = Ym[(size_t)(ao + kTriI[k]) * n + ao + kTriJ[k]];
    }
    const double ppt = pd[2] + pd[5] + pd[9], ypt = yd[2] + yd[5] + yd[9];
    const double ysp = yd[1] * pd[1] + yd[3] * pd[3] + yd[6] * pd[6];
    const double ypd = yd[2] * pd[2] + yd[5] * pd[5] + yd[9] * pd[9];
    const double ypo = yd[4] * pd[4] + yd[7] * pd[7] + yd[8] * pd[8];
    o[PAR_UPP] = ypt;
    o[PAR_GSP] = yss * ppt + ypt * pss - ysp;
    o[PAR_HSP] = -0.5 * (yss * ppt + ypt * pss) + 3.0 * ysp;
    o[PAR_GPP] = 0.75 * ypd - 0.25 * ypt * ppt + 1.5 * ypo;
    o[PAR_GP2] = 1.25 * (ypt * ppt - ypd) - 2.5 * ypo;
  }
}

// Diagonal core-Hamiltonian blocks: U on the diagonal plus the electron-core attractions
// collected from every tile the atom appears in (hcore.py:127-150), fixed summation order.
template <int G>
__global__ __launch_bounds__(256) void k_hcore_diag(pyseqm_layout_t L, const double* __restrict__ par,
                                                    const double* __restrict__ part, double* __restrict__ H) {
  const int a = blockIdx.x * (256 / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (a >= L.natoms) return;  // uniform over the lane group
  const double* p = par + (size_t)a * PYSEQM_NPAR;
  const int m = L.atom_mol[a];
  const int norb = L.mol_norb[m];
  double* Hm = H + L.mol_mat_off[m];
  const int ao = L.atom_ao[a];
  const int e0 = L.csr_off[a], e1 = L.csr_off[a + 1];
  if (L.atom_Z[a] > 1) {
    double acc[10];
    gather_slots_c<G>(part, L, e0, e1, lane, acc);
    acc[0] += p[PAR_USS];
    acc[2] += p[PAR_UPP];
    acc[5] += p[PAR_UPP];
    acc[9] += p[PAR_UPP];
    if (lane < 16) {
      const int mu = lane >> 2, nu = lane & 3;
      double v = acc[0];
#pragma unroll
      for (int k = 1; k < 10; ++k)
        if (k == tri_index(mu, nu)) v = acc[k];
      Hm[(size_t)(ao + mu) * norb + ao + nu] = v;
    }
  } else {
    double acc[1];
    gather_slots_c<G>(part, L, e0, e1, lane, acc);
    if (lane == 0) Hm[(size_t)ao * norb + ao] = acc[0] + p[PAR_USS];
  }
}

// Per-molecule sum of per-tile values (and optionally per-atom values): one block per molecule,
// threads stride over the molecule's tiles / atoms, fixed thread assignment, wave shuffle trees and
// a fixed-order sum of the wave totals.
template <int NT>
__global__ __launch_bounds__(NT) void k_mol_sum(pyseqm_layout_t L, const double* __restrict__ tile_v,
                                                const double* __restrict__ atom_v,
                                                const unsigned char* __restrict__ active, double* __restrict__ out) {
  __shared__ double wsum[NT / 64];
  const int m = blockIdx.x;
  if (active && !active[m]) return;
  const int tid = threadIdx.x;
  // eight independent partial sums per thread keep eight loads in flight (a single running sum
  // waits out one memory round trip per element); combined in a fixed order
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  auto add_range = [&](const double* v, int b, int e) {
    int t = b + tid;
    for (; t + 7 * NT < e; t += 8 * NT) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += v[t + u * NT];
    }
    for (; t < e; t += NT) acc[0] += v[t];
  };
  for (int kind = 0; kind < 3; ++kind) {
    const int* off = L.mol_tile_off + kind * (L.nmol + 1);
    add_range(tile_v, off[m], off[m + 1]);
  }
  if (atom_v) add_range(atom_v, L.mol_atom_off[m], L.mol_atom_off[m + 1]);
  double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  s = wave_sum(s);
  if ((tid & 63) == 0) wsum[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) r += wsum[w];
    out[m] = r;
  }
}

}  // namespace

namespace pyseqm {
void launch_mol_sum(const pyseqm_layout_t* L, const double* tile_v, const double* atom_v,
                    const unsigned char* active, double* out, hipStream_t s) {
  if (L->nmol <= 0) return;
  const long long nt = (long long)L->ntile[0] + L->ntile[1] + L->ntile[2];
  if (nt > 2048LL * L->nmol)
    k_mol_sum<1024><<<L->nmol, 1024, 0, s>>>(*L, tile_v, atom_v, active, out);
  else
    k_mol_sum<64><<<L->nmol, 64, 0, s>>>(*L, tile_v, atom_v, active, out);
}
}  // namespace pyseqm

namespace {
// Core-core repulsion of an explicit pair list (energy.py:91-134): one thread per pair, gam given.
__global__ void k_core_repulsion_pairs(long long npairs, int method, const int* __restrict__ idxi,
                                       const int* __restrict__ idxj, const double* __restrict__ rij,
                                       const double* __restrict__ gam, const double* __restrict__ par,
                                       double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const double* pi = par + (size_t)idxi[k] * PYSEQM_NPAR;
  const double* pj = par + (size_t)idxj[k] * PYSEQM_NPAR;
  out[k] = core_repulsion(method, (int)pi[PAR_Z], (int)pj[PAR_Z], rij[k], gam[k], pi, pj);
}
}  // namespace

extern "C" int pyseqm_core_repulsion_pairs(long long npairs, int method, const int* idxi, const int* idxj,
                                           const double* rij, const double* gam, const double* atom_par,
                                           double* enuc_pair, void* stream) {
  if (npairs < 0 || method < PYSEQM_MNDO || method > PYSEQM_PM3) return PYSEQM_EINVAL;
  if (npairs == 0) return PYSEQM_OK;
  k_core_repulsion_pairs<<<(unsigned)((npairs + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      npairs, method, idxi, idxj, rij, gam, atom_par, enuc_pair);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

namespace {
template <bool XL>
int launch_pair_forces(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                       const double* P, const double* D, double h, double* part, double* forces, void* stream) {
  if (!L || method < PYSEQM_MNDO || method > PYSEQM_PM3 || !(h > 0.0)) return PYSEQM_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  int base = 0;
  if (L->ntile[0])
    k_pair_forces<XX, XL><<<L->ntile[0], 64, 0, s>>>(*L, method, base, coords, atom_par, P, D, h, part);
  base += L->ntile[0];
  if (L->ntile[1])
    k_pair_forces<XH, XL><<<L->ntile[1], 64, 0, s>>>(*L, method, base, coords, atom_par, P, D, h, part);
  base += L->ntile[1];
  if (L->ntile[2])
    k_pair_forces<HH, XL><<<L->ntile[2], 64, 0, s>>>(*L, method, base, coords, atom_par, P, D, h, part);
  if (L->natoms) {
    if (lanes_per_atom(L) == 64)
      k_atom_forces<64><<<(L->natoms + 3) / 4, 256, 0, s>>>(*L, part, forces);
    else
      k_atom_forces<16><<<(L->natoms + 15) / 16, 256, 0, s>>>(*L, part, forces);
  }
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}
}  // namespace

extern "C" int pyseqm_pair_forces(const pyseqm_layout_t* L, int method, const double* coords,
                                  const double* atom_par, const double* P, double h, double* part, double* forces,
                                  void* stream) {
  return launch_pair_forces<false>(L, method, coords, atom_par, P, nullptr, h, part, forces, stream);
}

extern "C" int pyseqm_pair_forces_xl(const pyseqm_layout_t* L, int method, const double* coords,
                                     const double* atom_par, const double* D, const double* P, double h,
                                     double* part, double* forces, void* stream) {
  if (!D || !P) return PYSEQM_EINVAL;
  return launch_pair_forces<true>(L, method, coords, atom_par, P, D, h, part, forces, stream);
}

extern "C" long long pyseqm_param_grad_workspace(const pyseqm_layout_t* L) {
  if (!L || L->natoms < 0) return -1;
  return (long long)L->natoms * kParamGroup * 4 * PYSEQM_NPAR;
}

namespace {
template <int MODE>
int launch_param_grad(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                      const double* P, const double* Y, double hrel, double* rec, double* part, double* out,
                      void* stream) {
  if (!L || method < PYSEQM_MNDO || method > PYSEQM_PM3 || !(hrel > 0.0)) return PYSEQM_EINVAL;
  if (L->natoms == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_param_onecenter<<<(L->natoms + 127) / 128, 128, 0, s>>>(*L, P, Y, out);
  const int npp = pair_param_count(method);
  for (int p0 = 0; p0 < npp; p0 += kParamGroup) {
    const int np = npp - p0 < kParamGroup ? npp - p0 : kParamGroup;
    const long long nthr = (long long)L->natoms * kParamGroup * 4;
    k_param_records<<<(unsigned)((nthr + 255) / 256), 256, 0, s>>>(*L, atom_par, p0, np, hrel, rec);
    int base = 0;
    if (L->ntile[0])
      k_pair_param_grad<XX, MODE><<<L->ntile[0], 64, 0, s>>>(*L, method, base, coords, atom_par, P, Y, rec, p0, np,
                                                             hrel, part);
    base += L->ntile[0];
    if (L->ntile[1])
      k_pair_param_grad<XH, MODE><<<L->ntile[1], 64, 0, s>>>(*L, method, base, coords, atom_par, P, Y, rec, p0, np,
                                                             hrel, part);
    base += L->ntile[1];
    if (L->ntile[2])
      k_pair_param_grad<HH, MODE><<<L->ntile[2], 64, 0, s>>>(*L, method, base, coords, atom_par, P, Y, rec, p0, np,
                                                             hrel, part);
    if (lanes_per_atom(L) == 64)
      k_param_gather<64><<<(L->natoms + 3) / 4, 256, 0, s>>>(*L, part, p0, np, out);
    else
      k_param_gather<16><<<(L->natoms + 15) / 16, 256, 0, s>>>(*L, part, p0, np, out);
  }
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}
}  // namespace

extern "C" int pyseqm_param_grad(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                                 const double* P, double hrel, double* rec, double* part, double* out,
                                 void* stream) {
  return launch_param_grad<0>(L, method, coords, atom_par, P, nullptr, hrel, rec, part, out, stream);
}

extern "C" int pyseqm_fock_param_vjp(const pyseqm_layout_t* L, int method, const double* coords,
                                     const double* atom_par, const double* P, const double* Y, double hrel,
                                     double* rec, double* part, double* out, void* stream) {
  if (!Y) return PYSEQM_EINVAL;
  return launch_param_grad<2>(L, method, coords, atom_par, P, Y, hrel, rec, part, out, stream);
}

extern "C" int pyseqm_atom_setup(const pyseqm_layout_t* L, double* atom_par, void* stream) {
  if (!L || L->natoms < 0) return PYSEQM_EINVAL;
  if (L->natoms == 0) return PYSEQM_OK;
  hipStream_t s = (hipStream_t)stream;
  k_atom_setup<<<(L->natoms + 127) / 128, 128, 0, s>>>(*L, atom_par);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}

extern "C" int pyseqm_pair_integrals(const pyseqm_layout_t* L, int method, const double* coords,
                                     const double* atom_par, double* w_xx, double* w_xh, double* w_hh,
                                     double* H, double* part, double* enuc_tile, double* enuc_mol,
                                     double* enuc_pair, const int* pair_ref, void* stream) {
  if (!L || method < PYSEQM_MNDO || method > PYSEQM_PM3) return PYSEQM_EINVAL;
  if (enuc_pair && !pair_ref) return PYSEQM_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  int base = 0;
  if (L->ntile[0])
    k_pair_tiles<XX><<<L->ntile[0], 64, 0, s>>>(*L, method, base, coords, atom_par, w_xx, H, part, enuc_tile,
                                                enuc_pair, pair_ref);
  base += L->ntile[0];
  if (L->ntile[1])
    k_pair_tiles<XH><<<L->ntile[1], 64, 0, s>>>(*L, method, base, coords, atom_par, w_xh, H, part, enuc_tile,
                                                enuc_pair, pair_ref);
  base += L->ntile[1];
  if (L->ntile[2])
    k_pair_tiles<HH><<<L->ntile[2], 64, 0, s>>>(*L, method, base, coords, atom_par, w_hh, H, part, enuc_tile,
                                                enuc_pair, pair_ref);
  if (L->natoms) {
    if (lanes_per_atom(L) == 64)
      k_hcore_diag<64><<<(L->natoms + 3) / 4, 256, 0, s>>>(*L, atom_par, part, H);
    else
      k_hcore_diag<16><<<(L->natoms + 15) / 16, 256, 0, s>>>(*L, atom_par, part, H);
  }
  pyseqm::launch_mol_sum(L, enuc_tile, nullptr, nullptr, enuc_mol, s);
  return hipGetLastError() == hipSuccess ? PYSEQM_OK : PYSEQM_ELAUNCH;
}
