"""FP64 numpy restatement of the PYSEQM closed-shell sp SCF path (the oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The oracle works in the reference's own padded layout (``4*molsize`` orbitals
per molecule, 4x4 atom blocks) so that it checks the product's packed,
padding-free layout from the outside.  Paths are relative to the lanl/PYSEQM
repository root.

Pinning (what each part is checked against, tests/test_oracle_pins.py and
tests/test_oracle_golden.py):

* parse / parameters / nocc      -> reference goldens (methane, benzene),
                                    odd-electron error message golden.
* multipoles (dd, qq, rho0/1/2)  -> closed forms and the defining equations
                                    of rho1/rho2 (residual after 5 secant steps).
* local-frame integrals (ri)     -> an independent point-charge (Klopman-Ohno
                                    multipole) model evaluated directly.
* two-center tensor rotation     -> rotation invariance; the result does not
                                    depend on how the pi axes are completed.
* STO overlaps                   -> numerical quadrature of normalised STOs.
* one-center Fock terms          -> general NDDO formula with the one-center
                                    integral tensor.
* fock / energies / SCF          -> reference goldens for CH4 (MNDO, AM1, PM3)
                                    and CH4 + C6H6 (AM1): Etot/Eelec/Enuc.
* forces (Hellmann-Feynman, FD)  -> reference force goldens (CH4 x3 methods,
                                    CH4 + C6H6, methanal x3 AM1).
* density step (eigh / SP2)      -> idempotency, trace = 2 nocc, [F, P] = 0,
                                    SP2 -> projector of the same F.
* Pulay DIIS                     -> converges to the same fixed point as the
                                    adaptive mixer (energies equal within eps).
  The reference's batch-global DIIS counter is reproduced; iteration counts
  themselves are "parity unpinned" (no golden holds them).
"""
import math
import os

import numpy as np

# ---------------------------------------------------------------- constants
# seqm/seqm_functions/constants.py:4-23
EV = 27.21
A0 = 0.529167
EV_KCAL = 23.061
OVERLAP_CUTOFF = 40.0  # bohr, constants.py:23 (hcore.py:83 uses rij <= cutoff)

# Element tables for Z = 0..18, seqm/seqm_functions/constants.py:72-189.
TORE = np.array([0, 1, 0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3, 4, 5, 6, 7, 0], float)
QN = np.array([0, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3], int)
USSC = np.array([0, 1, 0, 1, 2, 2, 2, 2, 2, 2, 0, 1, 2, 2, 2, 2, 2, 2, 0], float)
UPPC = np.array([0, 0, 0, 0, 0, 1, 2, 3, 4, 5, 6, 0, 0, 1, 2, 3, 4, 5, 6], float)
GSSC = np.array([0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 0], float)
GSPC = np.array([0, 0, 0, 0, 0, 2, 4, 6, 8, 10, 0, 0, 0, 2, 4, 6, 8, 10, 0], float)
HSPC = np.array([0, 0, 0, 0, 0, -1, -2, -3, -4, -5, 0, 0, 0, -1, -2, -3, -4, -5, 0], float)
GP2C = np.array([0, 0, 0, 0, 0, 0, 1.5, 4.5, 6.5, 10, 0, 0, 0, 0, 1.5, 4.5, 6.5, 10, 0], float)
GPPC = np.array([0, 0, 0, 0, 0, 0, -0.5, -1.5, -0.5, 0, 0, 0, 0, 0, -0.5, -1.5, -0.5, 0, 0], float)
EHEAT_KCAL = np.array([0.0, 52.102, 0.0, 38.410, 76.960, 135.700, 170.890, 113.000, 59.559,
                       18.890, 0.0, 25.850, 35.000, 79.490, 108.390, 75.570, 66.400, 28.990, 0.0])
MASS = np.array([0.0, 1.00790, 4.00260, 6.94000, 9.01218, 10.81000, 12.01100, 14.00670, 15.99940,
                 18.99840, 20.17900, 22.98977, 24.30500, 26.98154, 28.08550, 30.97376, 32.06000,
                 35.45300, 39.94800])

PARAM_NAMES = ["U_ss", "U_pp", "zeta_s", "zeta_p", "beta_s", "beta_p", "g_ss", "g_sp", "g_pp",
               "g_p2", "h_sp", "alpha"] + [f"Gaussian{i}_{x}" for i in range(1, 5) for x in "KLM"]
N_GAUSS = {"MNDO": 0, "AM1": 4, "PM3": 2}  # seqm/basics.py:721-735

_PARAM_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "pyseqm_amd", "params")


def load_params(method, param_dir=None):
    """Per-element MOPAC parameters indexed by Z (seqm/seqm_functions/parameters.py:4-46), from the
    package's parameter files or the files under `param_dir` (seqm_parameters["parameter_file_dir"])."""
    if method not in N_GAUSS:
        raise ValueError(f"unsupported method {method}")
    table = {k: np.zeros(19) for k in PARAM_NAMES}
    with open(os.path.join(param_dir or _PARAM_DIR, f"parameters_{method}.csv")) as fh:
        hdr = fh.readline().strip().split(",")
        for line in fh:
            t = line.strip().split(",")
            z = int(t[0])
            for k in PARAM_NAMES:
                table[k][z] = float(t[hdr.index(k)])
    return table


# ------------------------------------------------------------------ parser
class Batch:
    pass


def parse(species, coords, method, charges=None, param_dir=None):
    """Pair list and per-atom data, seqm/basics.py:219-403 and Molecule.py:188-206."""
    species = np.asarray(species, dtype=np.int64)
    coords = np.asarray(coords, dtype=np.float64)
    nmol, m = species.shape
    bad = ~np.all(species[:, :-1] >= species[:, 1:], axis=1)
    if bad.any():
        rows = ", ".join(str(int(r)) for r in np.nonzero(bad)[0])
        raise ValueError(f"species must be non-increasing along each row, but row(s) {rows} not sorted.")
    b = Batch()
    b.method = method
    b.nmol, b.molsize = nmol, m
    real = species > 0
    flat = np.nonzero(real.reshape(-1))[0]
    b.Z = species.reshape(-1)[flat]
    b.atom_mol = flat // m
    b.atom_loc = flat % m
    b.natoms = len(flat)
    b.nHeavy = (species > 1).sum(1)
    b.nHydro = (species == 1).sum(1)
    nel = TORE[species].sum(1).astype(np.int64)
    if charges is not None:
        nel = nel - np.asarray(charges, dtype=np.int64).reshape(-1)
    if np.any(nel % 2 == 1):
        raise ValueError("RHF setting requires closed shell systems (even number of electrons)")
    b.nocc = nel // 2
    b.norb = 4 * b.nHeavy + b.nHydro
    b.maskd = b.atom_mol * m * m + b.atom_loc * (m + 1)
    iu, ju = np.triu_indices(m, 1)
    pm = np.repeat(np.arange(nmol), len(iu))
    pi = np.tile(iu, nmol)
    pj = np.tile(ju, nmol)
    keep = real[pm, pi] & real[pm, pj]
    pm, pi, pj = pm[keep], pi[keep], pj[keep]
    inv = np.full(nmol * m, -1, dtype=np.int64)
    inv[flat] = np.arange(b.natoms)
    b.idxi = inv[pm * m + pi]
    b.idxj = inv[pm * m + pj]
    b.ni = b.Z[b.idxi]
    b.nj = b.Z[b.idxj]
    d = coords[pm, pj] - coords[pm, pi]
    dist = np.sqrt((d * d).sum(1))
    b.xij = d / dist[:, None]
    b.rij = dist * (1.0 / A0)
    b.mask = pm * m * m + pi * m + pj
    b.pair_mol = pm
    b.species = species
    b.coords = coords
    b.prm = load_params(method, param_dir)
    return b


# -------------------------------------------------------------- multipoles
def dd_qq(qn, zs, zp):
    """Charge separations, seqm/seqm_functions/cal_par.py:11-28."""
    dd = (2.0 * qn + 1.0) * (4.0 * zs * zp) ** (qn + 0.5) / (zs + zp) ** (2.0 * qn + 2.0) / math.sqrt(3.0)
    qq = np.sqrt((4.0 * qn ** 2 + 6.0 * qn + 2.0) / 20.0) / zp
    return dd, qq


def _secant(target, f, x1):
    """Five fixed secant steps starting from x1 and x1+0.04 (cal_par.py:141-148, 226-240)."""
    x2 = x1 + 0.04
    for _ in range(5):
        f1 = f(x1)
        f2 = f(x2)
        ok = np.abs(f2 - f1) > 1.0e-16
        with np.errstate(divide="ignore", invalid="ignore"):
            x3 = np.where(ok, x1 + (x2 - x1) * (target - f1) / (f2 - f1), x2)
        x1, x2 = x2, x3
    return x2


def rho1_of(hsp_ev, D1):
    """rho1 = 1/(2 ad) with hsp = ad/2 - 1/(2 sqrt(4 D1^2 + 1/ad^2)), cal_par.py:106-160."""
    hsp = hsp_ev / EV
    d1 = (np.abs(hsp) / D1 ** 2) ** (1.0 / 3.0)
    d1 = np.where(hsp < 0.0, -d1, d1)
    f = lambda d: 0.5 * d - 0.5 / np.sqrt(4.0 * D1 ** 2 + 1.0 / d ** 2)
    return 0.5 / _secant(hsp, f, d1)


def rho2_of(hpp_ev, D2):
    """rho2 = 1/(2 aq), cal_par.py:192-246."""
    hpp = hpp_ev / EV
    q1 = (np.abs(hpp) / 3.0 / D2 ** 4) ** 0.2
    q1 = np.where(hpp < 0.0, -q1, q1)
    f = lambda q: (0.25 * q - 0.5 / np.sqrt(4.0 * D2 ** 2 + 1.0 / q ** 2)
                   + 0.25 / np.sqrt(8.0 * D2 ** 2 + 1.0 / q ** 2))
    return 0.5 / _secant(hpp, f, q1)


def atom_multipoles(Z, prm):
    """dd, qq, rho0, rho1, rho2 per atom, two_elec_two_center_int.py:117-147, 240-243."""
    gss, gpp, gp2, hsp = prm["g_ss"][Z], prm["g_pp"][Z], prm["g_p2"][Z], prm["h_sp"][Z]
    zs, zp = prm["zeta_s"][Z], prm["zeta_p"][Z]
    heavy = Z > 2
    n = len(Z)
    dd, qq, r1, r2 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    qn = QN[Z].astype(float)
    dd[heavy], qq[heavy] = dd_qq(qn[heavy], zs[heavy], zp[heavy])
    rho0 = 0.5 * EV / gss
    hpp = np.maximum(0.5 * (gpp - gp2), 0.1)
    r1[heavy] = rho1_of(hsp[heavy], dd[heavy])
    r2[heavy] = rho2_of(hpp[heavy], qq[heavy])
    return dd, qq, rho0, r1, r2


# ----------------------------------------------------- local-frame integrals
def local_ri_hh(r, r0a, r0b):
    """(ss|ss) for H-H, two_elec_two_center_int_local_frame.py:77."""
    return EV / np.sqrt(r ** 2 + (r0a + r0b) ** 2)


def local_ri_xh(r, da, qa0, r0a, r1a, r2a, r0b):
    """4 local integrals heavy-H, two_elec_two_center_int_local_frame.py:90-107."""
    e1, e2 = EV / 2.0, EV / 4.0
    qa = 2.0 * qa0
    aee = (r0a + r0b) ** 2
    ade = (r1a + r0b) ** 2
    aqe = (r2a + r0b) ** 2
    ee = EV / np.sqrt(r ** 2 + aee)
    cen = e1 / np.sqrt(r ** 2 + aqe)
    out = np.empty((len(r), 4))
    out[:, 0] = ee
    out[:, 1] = e1 / np.sqrt((r + da) ** 2 + ade) - e1 / np.sqrt((r - da) ** 2 + ade)
    out[:, 2] = ee + e2 / np.sqrt((r + qa) ** 2 + aqe) + e2 / np.sqrt((r - qa) ** 2 + aqe) - cen
    out[:, 3] = ee + e1 / np.sqrt(r ** 2 + qa ** 2 + aqe) - cen
    return out


def local_ri_xx(r, da, db, qa0, qb0, r0a, r0b, r1a, r1b, r2a, r2b):
    """22 local integrals heavy-heavy (MOPAC repp), two_elec_two_center_int_local_frame.py:126-262."""
    e1, e2, e3, e4 = EV / 2.0, EV / 4.0, EV / 8.0, EV / 16.0
    s = np.sqrt
    qa, qb = 2.0 * qa0, 2.0 * qb0
    aee = (r0a + r0b) ** 2
    ade = (r1a + r0b) ** 2
    aqe = (r2a + r0b) ** 2
    aed = (r0a + r1b) ** 2
    aeq = (r0a + r2b) ** 2
    axx = (r1a + r1b) ** 2
    adq = (r1a + r2b) ** 2
    aqd = (r2a + r1b) ** 2
    aqq = (r2a + r2b) ** 2
    ee = EV / s(r ** 2 + aee)
    dze = -e1 / s((r + da) ** 2 + ade) + e1 / s((r - da) ** 2 + ade)
    qzze = e2 / s((r - qa) ** 2 + aqe) + e2 / s((r + qa) ** 2 + aqe) - e1 / s(r ** 2 + aqe)
    qxxe = e1 / s(r ** 2 + qa ** 2 + aqe) - e1 / s(r ** 2 + aqe)
    edz = -e1 / s((r - db) ** 2 + aed) + e1 / s((r + db) ** 2 + aed)
    eqzz = e2 / s((r - qb) ** 2 + aeq) + e2 / s((r + qb) ** 2 + aeq) - e1 / s(r ** 2 + aeq)
    eqxx = e1 / s(r ** 2 + qb ** 2 + aeq) - e1 / s(r ** 2 + aeq)
    t20 = e2 / s((r + da) ** 2 + adq)
    t22 = e2 / s((r - da) ** 2 + adq)
    t24 = e2 / s((r - db) ** 2 + aqd)
    t26 = e2 / s((r + db) ** 2 + aqd)
    t36 = e2 / s(r ** 2 + aqq)
    t39 = e2 / s(r ** 2 + qa ** 2 + aqq)
    t40 = e2 / s(r ** 2 + qb ** 2 + aqq)
    t42 = e3 / s((r - qb) ** 2 + aqq)
    t44 = e3 / s((r + qb) ** 2 + aqq)
    t46 = e3 / s((r + qa) ** 2 + aqq)
    t48 = e3 / s((r - qa) ** 2 + aqq)
    ri = np.empty((len(r), 22))
    ri[:, 0] = ee
    ri[:, 1] = -dze
    ri[:, 2] = ee + qzze
    ri[:, 3] = ee + qxxe
    ri[:, 4] = -edz
    ri[:, 5] = (e2 / s((r + da - db) ** 2 + axx) + e2 / s((r - da + db) ** 2 + axx)
                - e2 / s((r - da - db) ** 2 + axx) - e2 / s((r + da + db) ** 2 + axx))
    ri[:, 6] = e1 / s(r ** 2 + (da - db) ** 2 + axx) - e1 / s(r ** 2 + (da + db) ** 2 + axx)
    ri[:, 7] = (-edz + e3 / s((r + qa - db) ** 2 + aqd) - e3 / s((r + qa + db) ** 2 + aqd)
                + e3 / s((r - qa - db) ** 2 + aqd) - e3 / s((r - qa + db) ** 2 + aqd) - t24 + t26)
    ri[:, 8] = (-edz - t24 + e2 / s((r - db) ** 2 + qa ** 2 + aqd) + t26
                - e2 / s((r + db) ** 2 + qa ** 2 + aqd))
    ri[:, 9] = (e2 / s((qa0 - db) ** 2 + (r + qa0) ** 2 + aqd) - e2 / s((qa0 - db) ** 2 + (r - qa0) ** 2 + aqd)
                - e2 / s((qa0 + db) ** 2 + (r + qa0) ** 2 + aqd)
                + e2 / s((qa0 + db) ** 2 + (r - qa0) ** 2 + aqd))
    ri[:, 10] = ee + eqzz
    ri[:, 11] = ee + eqxx
    ri[:, 12] = (-dze + e3 / s((r + da - qb) ** 2 + adq) - e3 / s((r - da - qb) ** 2 + adq)
                 + e3 / s((r + da + qb) ** 2 + adq) - e3 / s((r - da + qb) ** 2 + adq) + t22 - t20)
    ri[:, 13] = (-dze - t20 + e2 / s((r + da) ** 2 + qb ** 2 + adq) + t22
                 - e2 / s((r - da) ** 2 + qb ** 2 + adq))
    ri[:, 14] = (e2 / s((da - qb0) ** 2 + (r - qb0) ** 2 + adq) - e2 / s((da - qb0) ** 2 + (r + qb0) ** 2 + adq)
                 - e2 / s((da + qb0) ** 2 + (r - qb0) ** 2 + adq)
                 + e2 / s((da + qb0) ** 2 + (r + qb0) ** 2 + adq))
    ri[:, 15] = (ee + eqzz + qzze + e4 / s((r + qa - qb) ** 2 + aqq) + e4 / s((r + qa + qb) ** 2 + aqq)
                 + e4 / s((r - qa - qb) ** 2 + aqq) + e4 / s((r - qa + qb) ** 2 + aqq)
                 - t48 - t46 - t42 - t44 + t36)
    ri[:, 16] = (ee + eqzz + qxxe + e3 / s((r - qb) ** 2 + qa ** 2 + aqq)
                 + e3 / s((r + qb) ** 2 + qa ** 2 + aqq) - t42 - t44 - t39 + t36)
    ri[:, 17] = (ee + eqxx + qzze + e3 / s((r + qa) ** 2 + qb ** 2 + aqq)
                 + e3 / s((r - qa) ** 2 + qb ** 2 + aqq) - t46 - t48 - t40 + t36)
    qxxqxx = e3 / s(r ** 2 + (qa - qb) ** 2 + aqq) + e3 / s(r ** 2 + (qa + qb) ** 2 + aqq) - t39 - t40 + t36
    ri[:, 18] = ee + eqxx + qxxe + qxxqxx
    a1, b1 = qa0, qb0
    ri[:, 19] = (e3 / s((r + a1 - b1) ** 2 + (a1 - b1) ** 2 + aqq) - e3 / s((r + a1 + b1) ** 2 + (a1 - b1) ** 2 + aqq)
                 - e3 / s((r - a1 - b1) ** 2 + (a1 - b1) ** 2 + aqq) + e3 / s((r - a1 + b1) ** 2 + (a1 - b1) ** 2 + aqq)
                 - e3 / s((r + a1 - b1) ** 2 + (a1 + b1) ** 2 + aqq) + e3 / s((r + a1 + b1) ** 2 + (a1 + b1) ** 2 + aqq)
                 + e3 / s((r - a1 - b1) ** 2 + (a1 + b1) ** 2 + aqq) - e3 / s((r - a1 + b1) ** 2 + (a1 + b1) ** 2 + aqq))
    qxxqyy = e2 / s(r ** 2 + qa ** 2 + qb ** 2 + aqq) - t39 - t40 + t36
    ri[:, 20] = ee + eqxx + qxxe + qxxqyy
    ri[:, 21] = 0.5 * (qxxqxx - qxxqyy)
    return ri


# local basis order: 0 = s, 1 = p-sigma, 2 = p-pi, 3 = p-pi'
def local_tensor_from_ri(ri):
    """(ab|cd) in the local frame from the 22 MOPAC integrals.

    Storage order of ri follows two_elec_two_center_int_local_frame.py:43-49
    ((SS/SS)=1 ... (P*P/P*P)=22); the tensor is symmetric in (a,b) and (c,d).
    """
    n = ri.shape[0]
    L = np.zeros((n, 4, 4, 4, 4))

    def put(a, b, c, d, v):
        for (p, q) in {(a, b), (b, a)}:
            for (u, w) in {(c, d), (d, c)}:
                L[:, p, q, u, w] = v

    S, O = 0, 1
    for P in (2, 3):
        Q = 5 - P
        put(P, P, S, S, ri[:, 3])
        put(S, P, S, P, ri[:, 6])
        put(P, P, S, O, ri[:, 8])
        put(O, P, S, P, ri[:, 9])
        put(S, S, P, P, ri[:, 11])
        put(S, O, P, P, ri[:, 13])
        put(S, P, O, P, ri[:, 14])
        put(P, P, O, O, ri[:, 16])
        put(O, O, P, P, ri[:, 17])
        put(P, P, P, P, ri[:, 18])
        put(O, P, O, P, ri[:, 19])
        put(P, P, Q, Q, ri[:, 20])
    put(S, S, S, S, ri[:, 0])
    put(S, O, S, S, ri[:, 1])
    put(O, O, S, S, ri[:, 2])
    put(S, S, S, O, ri[:, 4])
    put(S, O, S, O, ri[:, 5])
    put(O, O, S, O, ri[:, 7])
    put(S, S, O, O, ri[:, 10])
    put(S, O, O, O, ri[:, 12])
    put(O, O, O, O, ri[:, 15])
    put(2, 3, 2, 3, ri[:, 21])
    return L


def frame_from_axis(v):
    """Rotation R with R @ v = e_x, two_elec_two_center_int.py:1576-1631 (quaternion form).

    Row 0 of R is v; rows 1 and 2 complete an orthonormal frame.
    """
    n = v.shape[0]
    q = np.zeros((n, 4))
    q[:, 1] = v[:, 2]
    q[:, 2] = -v[:, 1]
    q[:, 3] = 1.0 + v[:, 0]
    flip = np.abs(q[:, 3]) < 1.0e-7
    q[flip] = np.array([0.0, 0.0, 1.0, 0.0])
    q /= np.sqrt((q * q).sum(1))[:, None]
    qy, qz, qw = q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    R[:, 0, 1] = -2 * (qz * qw)
    R[:, 0, 2] = 2 * (qy * qw)
    R[:, 1, 0] = 2 * (qz * qw)
    R[:, 1, 1] = 1 - 2 * (qz * qz)
    R[:, 1, 2] = 2 * (qy * qz)
    R[:, 2, 0] = -2 * (qy * qw)
    R[:, 2, 1] = 2 * (qy * qz)
    R[:, 2, 2] = 1 - 2 * (qy * qy)
    return R


def rotate_tensor(L, R):
    """Molecular-frame (mu nu|la si) = sum U_mu,a U_nu,b U_la,c U_si,d (ab|cd).

    Molecular p_m = sum_a R[a, m] * local_a (two_elec_two_center_int.py:1384-1574
    writes the same transformation out term by term).
    """
    n = L.shape[0]
    U = np.zeros((n, 4, 4))
    U[:, 0, 0] = 1.0
    U[:, 1:, 1:] = np.transpose(R, (0, 2, 1))
    return np.einsum("nia,njb,nkc,nld,nabcd->nijkl", U, U, U, U, L, optimize=True)


TRIL = [(k, l) for k in range(4) for l in range(k + 1)]  # (kk, ll) order of w rows/cols


def tensor_to_w(T):
    """(npairs,4,4,4,4) -> reference compact (npairs,10,10) (two_elec_two_center_int.py:1414-1421)."""
    w = np.empty((T.shape[0], 10, 10))
    for a, (k, l) in enumerate(TRIL):
        for b, (m_, n_) in enumerate(TRIL):
            w[:, a, b] = T[:, k, l, m_, n_]
    return w


def two_center_tensor(b, mp=None):
    """Full (npairs,4,4,4,4) molecular-frame two-electron integrals plus ri/riXH.

    Restates two_elec_two_center_int.py:98-283 (+ rotate 286-372, 1301-1306) for sp atoms.
    """
    if mp is None:
        mp = atom_multipoles(b.Z, b.prm)
    dd, qq, rho0, rho1, rho2 = mp
    i, j = b.idxi, b.idxj
    npairs = len(i)
    T = np.zeros((npairs, 4, 4, 4, 4))
    L = np.zeros((npairs, 4, 4, 4, 4))
    HH = (b.ni == 1) & (b.nj == 1)
    XH = (b.ni > 1) & (b.nj == 1)
    XX = (b.ni > 1) & (b.nj > 1)
    r = b.rij
    L[HH, 0, 0, 0, 0] = local_ri_hh(r[HH], rho0[i[HH]], rho0[j[HH]])
    rxh = local_ri_xh(r[XH], dd[i[XH]], qq[i[XH]], rho0[i[XH]], rho1[i[XH]], rho2[i[XH]], rho0[j[XH]])
    LXH = np.zeros((XH.sum(), 4, 4, 4, 4))
    LXH[:, 0, 0, 0, 0] = rxh[:, 0]
    LXH[:, 0, 1, 0, 0] = LXH[:, 1, 0, 0, 0] = rxh[:, 1]
    LXH[:, 1, 1, 0, 0] = rxh[:, 2]
    LXH[:, 2, 2, 0, 0] = LXH[:, 3, 3, 0, 0] = rxh[:, 3]
    L[XH] = LXH
    rxx = local_ri_xx(r[XX], dd[i[XX]], dd[j[XX]], qq[i[XX]], qq[j[XX]], rho0[i[XX]], rho0[j[XX]],
                      rho1[i[XX]], rho1[j[XX]], rho2[i[XX]], rho2[j[XX]])
    L[XX] = local_tensor_from_ri(rxx)
    R = frame_from_axis(-b.xij)
    T = rotate_tensor(L, R)
    return T, rxh, rxx


# ------------------------------------------------------------- STO overlap
def aint(x, jcall):
    """A_k(x) = int_1^inf t^k exp(-x t) dt, k = 0..12 (diat_overlap_PM6_SP.py:464-519)."""
    x = np.where(x != 0, x, np.inf)
    a = np.zeros((len(x), 13))
    a[:, 0] = np.exp(-x) / x
    for k in range(1, 13):
        need = jcall >= k if k >= 3 else np.ones(len(x), bool)
        v = a[:, 0] + k * a[:, k - 1] / x
        a[:, k] = np.where(need, v, 0.0)
    return a


def bint(x):
    """B_k(x) = int_-1^1 t^k exp(-x t) dt, k = 0..12 (diat_overlap_PM6_SP.py:522-670).

    Recurrence for |x| > 0.5, the reference's truncated series for 1e-6 < |x| <= 0.5,
    and the x = 0 limits below that.
    """
    n = len(x)
    b = np.zeros((n, 13))
    for k in range(13):
        b[:, k] = 2.0 / (k + 1) if k % 2 == 0 else 0.0
    big = np.abs(x) > 0.5
    xb = x[big]
    if big.any():
        tx = np.exp(xb) / xb
        tmx = -np.exp(-xb) / xb
        prev = tx + tmx
        b[big, 0] = prev
        for k in range(1, 13):
            sgn = -1.0 if k % 2 == 1 else 1.0
            prev = sgn * tx + tmx + k * prev / xb
            b[big, k] = prev
    small = (np.abs(x) <= 0.5) & (np.abs(x) > 1.0e-6)
    xs = x[small]
    if small.any():
        even_den = {0: (3.0, 60.0, 2520.0), 2: (5.0, 84.0, 3240.0), 4: (7.0, 108.0, 3960.0),
                    6: (9.0, 132.0, 4680.0), 8: (11.0, 156.0, 5400.0), 10: (13.0, 180.0, 6120.0),
                    12: (15.0, 204.0, 6840.0)}
        for k, (d2, d4, d6) in even_den.items():
            b[small, k] = 2.0 / (k + 1) + xs ** 2 / d2 + xs ** 4 / d4 + xs ** 6 / d6
        odd_den = {1: (15.0, 420.0), 3: (21.0, 540.0), 5: (27.0, 660.0), 7: (33.0, 780.0),
                   9: (39.0, 900.0), 11: (45.0, 1020.0)}
        for k, (d3, d5) in odd_den.items():
            b[small, k] = -2.0 / (k + 2) * xs - xs ** 3 / d3 - xs ** 5 / d5
    return b


def _jcall(qni, qnj):
    jc = np.zeros(len(qni), dtype=np.int64)
    jc[(qni == 1) & (qnj == 1)] = 2
    jc[(qni == 2) & (qnj == 1)] = 3
    jc[(qni == 2) & (qnj == 2)] = 4
    jc[(qni == 3) & (qnj == 1)] = 431
    jc[(qni == 3) & (qnj == 2)] = 5
    jc[(qni == 3) & (qnj == 3)] = 6
    return jc


def overlap_local(ni, nj, rij, za, zb):
    """S111, S211, S121, S221, S222 in the diatomic frame (diat_overlap_PM6_SP.py:55-345).

    za, zb: (n, 2) zeta_s / zeta_p of atoms i and j (Z_i >= Z_j).
    """
    jc = _jcall(QN[ni], QN[nj])
    if np.any(jc == 0):
        raise ValueError("\nError from diat.py, overlap matrix\nSome elements are not supported yet")
    n = len(rij)
    r = rij

    def AB(z1, z2):
        return aint(0.5 * r * (z1 + z2), jc), bint(0.5 * r * (z1 - z2))

    A1, B1 = AB(za[:, 0], zb[:, 0])
    A2, B2 = AB(za[:, 1], zb[:, 0])
    A3, B3 = AB(za[:, 0], zb[:, 1])
    A4, B4 = AB(za[:, 1], zb[:, 1])
    S = np.zeros((5, n))
    s3 = math.sqrt(3.0)
    c = jc == 2
    if c.any():
        S[0, c] = ((za[c, 0] * zb[c, 0] * r[c] ** 2) ** 1.5
                   * (A1[c, 2] * B1[c, 0] - B1[c, 2] * A1[c, 0]) / 4.0)
    c = jc == 3
    if c.any():
        S[0, c] = (zb[c, 0] ** 1.5 * za[c, 0] ** 2.5 * r[c] ** 4
                   * (A1[c, 3] * B1[c, 0] - B1[c, 3] * A1[c, 0] + A1[c, 2] * B1[c, 1] - B1[c, 2] * A1[c, 1])
                   / (s3 * 8.0))
        S[1, c] = (zb[c, 0] ** 1.5 * za[c, 1] ** 2.5 * r[c] ** 4
                   * (A2[c, 2] * B2[c, 0] - B2[c, 2] * A2[c, 0] + A2[c, 3] * B2[c, 1] - B2[c, 3] * A2[c, 1])
                   / 8.0)
    c = jc == 4
    if c.any():
        r5 = r[c] ** 5
        S[0, c] = ((zb[c, 0] * za[c, 0]) ** 2.5 * r5
                   * (A1[c, 4] * B1[c, 0] + B1[c, 4] * A1[c, 0] - 2.0 * A1[c, 2] * B1[c, 2]) / 48.0)
        S[1, c] = ((zb[c, 0] * za[c, 1]) ** 2.5 * r5
                   * (A2[c, 3] * (B2[c, 0] - B2[c, 2]) - A2[c, 1] * (B2[c, 2] - B2[c, 4])
                      + B2[c, 3] * (A2[c, 0] - A2[c, 2]) - B2[c, 1] * (A2[c, 2] - A2[c, 4]))
                   / (16.0 * s3))
        S[2, c] = ((zb[c, 1] * za[c, 0]) ** 2.5 * r5
                   * (A3[c, 3] * (B3[c, 0] - B3[c, 2]) - A3[c, 1] * (B3[c, 2] - B3[c, 4])
                      - B3[c, 3] * (A3[c, 0] - A3[c, 2]) + B3[c, 1] * (A3[c, 2] - A3[c, 4]))
                   / (16.0 * s3))
        w = (zb[c, 1] * za[c, 1]) ** 2.5 * r5 / 16.0
        S[3, c] = -w * (B4[c, 2] * (A4[c, 4] + A4[c, 0]) - A4[c, 2] * (B4[c, 4] + B4[c, 0]))
        S[4, c] = 0.5 * w * (A4[c, 4] * (B4[c, 0] - B4[c, 2]) - B4[c, 4] * (A4[c, 0] - A4[c, 2])
                             - A4[c, 2] * B4[c, 0] + B4[c, 2] * A4[c, 0])
    c = jc == 431
    if c.any():
        r5 = r[c] ** 5
        S[0, c] = (zb[c, 0] ** 1.5 * za[c, 0] ** 3.5 * r5
                   * (A1[c, 4] * B1[c, 0] + 2 * B1[c, 1] * A1[c, 3] - 2.0 * A1[c, 1] * B1[c, 3]
                      - B1[c, 4] * A1[c, 0]) / (math.sqrt(10.0) * 24.0))
        S[1, c] = (zb[c, 0] ** 1.5 * za[c, 1] ** 3.5 * r5
                   * (A2[c, 3] * (B2[c, 0] + B2[c, 2]) - A2[c, 1] * (B2[c, 4] + B2[c, 2])
                      + B2[c, 1] * (A2[c, 2] + A2[c, 4]) - B2[c, 3] * (A2[c, 2] + A2[c, 0]))
                   / (8.0 * math.sqrt(30.0)))
    c = jc == 5
    if c.any():
        r6 = r[c] ** 6
        S[0, c] = (zb[c, 0] ** 2.5 * za[c, 0] ** 3.5 * r6
                   * (A1[c, 5] * B1[c, 0] + B1[c, 1] * A1[c, 4] - 2 * B1[c, 2] * A1[c, 3]
                      - 2.0 * A1[c, 2] * B1[c, 3] + B1[c, 4] * A1[c, 1] + B1[c, 5] * A1[c, 0])
                   / (math.sqrt(30.0) * 48.0))
        S[1, c] = (zb[c, 0] ** 2.5 * za[c, 1] ** 3.5 * r6
                   * (A2[c, 4] * B2[c, 0] + B2[c, 1] * A2[c, 5] - 2.0 * B2[c, 3] * A2[c, 3]
                      - 2.0 * A2[c, 2] * B2[c, 2] + A2[c, 1] * B2[c, 5] + A2[c, 0] * B2[c, 4])
                   / (48.0 * math.sqrt(10.0)))
        S[2, c] = (zb[c, 1] ** 2.5 * za[c, 0] ** 3.5 * r6
                   * ((A3[c, 4] * B3[c, 0] - A3[c, 5] * B3[c, 1])
                      + 2.0 * (A3[c, 3] * B3[c, 1] - A3[c, 4] * B3[c, 2])
                      - 2.0 * (A3[c, 1] * B3[c, 3] - A3[c, 2] * B3[c, 4])
                      - (A3[c, 0] * B3[c, 4] - A3[c, 1] * B3[c, 5]))
                   / (48.0 * math.sqrt(10.0)))
        S[3, c] = (zb[c, 1] ** 2.5 * za[c, 1] ** 3.5 * r6
                   * ((A4[c, 3] * B4[c, 0] - A4[c, 5] * B4[c, 2])
                      + (A4[c, 2] * B4[c, 1] - A4[c, 4] * B4[c, 3])
                      - (A4[c, 1] * B4[c, 2] - A4[c, 3] * B4[c, 4])
                      - (A4[c, 0] * B4[c, 3] - A4[c, 2] * B4[c, 5]))
                   / (16.0 * math.sqrt(30.0)))
        S[4, c] = (zb[c, 1] ** 2.5 * za[c, 1] ** 3.5 * r6
                   * ((A4[c, 5] - A4[c, 3]) * (B4[c, 0] - B4[c, 2])
                      + (A4[c, 4] - A4[c, 2]) * (B4[c, 1] - B4[c, 3])
                      - (A4[c, 3] - A4[c, 1]) * (B4[c, 2] - B4[c, 4])
                      - (A4[c, 2] - A4[c, 0]) * (B4[c, 3] - B4[c, 5]))
                   / (32.0 * math.sqrt(30.0)))
    c = jc == 6
    if c.any():
        r7 = r[c] ** 7
        S[0, c] = ((zb[c, 0] * za[c, 0]) ** 3.5 * r7
                   * (A1[c, 6] * B1[c, 0] - 3.0 * B1[c, 2] * A1[c, 4] + 3.0 * A1[c, 2] * B1[c, 4]
                      - A1[c, 0] * B1[c, 6]) / 1440.0)
        S[1, c] = ((zb[c, 0] * za[c, 1]) ** 3.5 * r7
                   * ((A2[c, 5] * B2[c, 0] + A2[c, 6] * B2[c, 1])
                      + (-A2[c, 4] * B2[c, 1] - A2[c, 5] * B2[c, 2])
                      - 2.0 * (A2[c, 3] * B2[c, 2] + A2[c, 4] * B2[c, 3])
                      - 2.0 * (-A2[c, 2] * B2[c, 3] - A2[c, 3] * B2[c, 4])
                      + (A2[c, 1] * B2[c, 4] + A2[c, 2] * B2[c, 5])
                      + (-A2[c, 0] * B2[c, 5] - A2[c, 1] * B2[c, 6])) / (480.0 * s3))
        S[2, c] = ((zb[c, 1] * za[c, 0]) ** 3.5 * r7
                   * ((A3[c, 5] * B3[c, 0] - A3[c, 6] * B3[c, 1])
                      + (A3[c, 4] * B3[c, 1] - A3[c, 5] * B3[c, 2])
                      - 2.0 * (A3[c, 3] * B3[c, 2] - A3[c, 4] * B3[c, 3])
                      - 2.0 * (A3[c, 2] * B3[c, 3] - A3[c, 3] * B3[c, 4])
                      + (A3[c, 1] * B3[c, 4] - A3[c, 2] * B3[c, 5])
                      + (A3[c, 0] * B3[c, 5] - A3[c, 1] * B3[c, 6])) / (480.0 * s3))
        S[3, c] = (zb[c, 1] ** 3.5 * za[c, 1] ** 3.5 * r7
                   * ((A4[c, 4] * B4[c, 0] - A4[c, 6] * B4[c, 2])
                      - 2.0 * (A4[c, 2] * B4[c, 2] - A4[c, 4] * B4[c, 4])
                      + (A4[c, 0] * B4[c, 4] - A4[c, 2] * B4[c, 6])) / 480.0)
        S[4, c] = (zb[c, 1] ** 3.5 * za[c, 1] ** 3.5 * r7
                   * ((A4[c, 6] - A4[c, 4]) * (B4[c, 0] - B4[c, 2])
                      - 2.0 * (A4[c, 4] - A4[c, 2]) * (B4[c, 2] - B4[c, 4])
                      + (A4[c, 2] - A4[c, 0]) * (B4[c, 4] - B4[c, 6])) / 960.0)
    return S


def overlap_blocks(b):
    """4x4 overlap <mu(i)|nu(j)> per pair in the molecular frame (diat_overlap_PM6_SP.py:369-385).

    <p_m|s> = S211 x_m, <s|p_n> = -S121 x_n, <p_m|p_n> = S222 delta_mn - (S221 + S222) x_m x_n,
    x = unit vector from atom i to atom j.
    """
    prm = b.prm
    zeta = np.stack([prm["zeta_s"][b.Z], prm["zeta_p"][b.Z]], 1)
    di = np.zeros((len(b.rij), 4, 4))
    ok = b.rij <= OVERLAP_CUTOFF
    if not ok.any():
        return di
    S = overlap_local(b.ni[ok], b.nj[ok], b.rij[ok], zeta[b.idxi[ok]], zeta[b.idxj[ok]])
    Rp = frame_from_axis(b.xij[ok])   # row 0 equals x except in the antipodal-flip case
    x = Rp[:, 0, :]
    d = np.zeros((ok.sum(), 4, 4))
    d[:, 0, 0] = S[0]
    d[:, 1:, 0] = S[1][:, None] * x
    d[:, 0, 1:] = -S[2][:, None] * x
    diag = np.stack([-S[3], S[4], S[4]], 1)
    d[:, 1:, 1:] = np.einsum("nai,na,nak->nik", Rp, diag, Rp)
    di[ok] = d
    return di


# ------------------------------------------------------------------ hcore
def hcore(b):
    """Symmetric padded core Hamiltonian + pair tensors (seqm/seqm_functions/hcore.py:9-179).

    Returns H (nmol, 4m, 4m) full symmetric, T (npairs,4,4,4,4), di (npairs,4,4) and the
    per-pair nuclear-repulsion gamma = (ss|ss).
    """
    prm = b.prm
    m = b.molsize
    mp = atom_multipoles(b.Z, prm)
    T, _, _ = two_center_tensor(b, mp)
    di = overlap_blocks(b)
    H = np.zeros((b.nmol, 4 * m, 4 * m))
    mol, loc = b.atom_mol, b.atom_loc
    for a in range(b.natoms):
        o = 4 * loc[a]
        H[mol[a], o, o] = prm["U_ss"][b.Z[a]]
        for k in range(1, 4):
            H[mol[a], o + k, o + k] = prm["U_pp"][b.Z[a]]
    # electron-core attraction: -Z_B (mu nu | s_B s_B) on atom A and vice versa (hcore.py:147-150)
    e1b = -TORE[b.nj][:, None, None] * T[:, :, :, 0, 0]
    e2a = -TORE[b.ni][:, None, None] * T[:, 0, 0, :, :]
    beta = np.stack([prm["beta_s"][b.Z]] + [prm["beta_p"][b.Z]] * 3, 1)
    bsum = 0.5 * (beta[b.idxi][:, :, None] + beta[b.idxj][:, None, :])
    hab = di * bsum  # hcore.py:155-173
    li, lj, pm = b.atom_loc[b.idxi], b.atom_loc[b.idxj], b.pair_mol
    Hb = H.reshape(b.nmol, m, 4, m, 4).transpose(0, 1, 3, 2, 4)  # view [mol, i, j, mu, nu]
    np.add.at(Hb, (pm, li, li), e1b)
    np.add.at(Hb, (pm, lj, lj), e2a)
    Hb[pm, li, lj] = hab
    Hb[pm, lj, li] = np.transpose(hab, (0, 2, 1))
    # H atoms carry only the s orbital: zero their p rows/cols
    for a in np.nonzero(b.Z == 1)[0]:
        o = 4 * loc[a]
        H[mol[a], o + 1:o + 4, :] = 0.0
        H[mol[a], :, o + 1:o + 4] = 0.0
    gam = T[:, 0, 0, 0, 0]
    return H, T, di, gam, mp


# ------------------------------------------------------------------ fock
def one_center_tensor(Z, prm):
    """One-center (ab|cd) from gss, gsp, gpp, gp2, hsp; hpp = (gpp - gp2)/2 (fock.py:196-224)."""
    n = len(Z)
    G = np.zeros((n, 4, 4, 4, 4))
    gss, gsp, gpp, gp2, hsp = (prm[k][Z] for k in ("g_ss", "g_sp", "g_pp", "g_p2", "h_sp"))
    hpp = 0.5 * (gpp - gp2)
    G[:, 0, 0, 0, 0] = gss
    for p in (1, 2, 3):
        G[:, 0, 0, p, p] = G[:, p, p, 0, 0] = gsp
        for (a, bb, c, d) in ((0, p, 0, p), (0, p, p, 0), (p, 0, 0, p), (p, 0, p, 0)):
            G[:, a, bb, c, d] = hsp
        G[:, p, p, p, p] = gpp
        for q in (1, 2, 3):
            if q != p:
                G[:, p, p, q, q] = gp2
                for (a, bb, c, d) in ((p, q, p, q), (p, q, q, p)):
                    G[:, a, bb, c, d] = hpp
    return G


def fock(b, P, H, T, G1=None):
    """F = H + G(P) for the closed-shell NDDO Hamiltonian (seqm/seqm_functions/fock.py:132-347).

    G_mu nu = sum_{la si} P_la si [(mu nu|la si) - 1/2 (mu la|nu si)] with one-center and
    two-center integral tensors; the two-center exchange only couples mu, la on different atoms.
    """
    if G1 is None:
        G1 = one_center_tensor(b.Z, b.prm)
    m = b.molsize
    F = H.copy()
    Pb = P.reshape(b.nmol, m, 4, m, 4).transpose(0, 1, 3, 2, 4)  # [mol, i, j, mu, nu]
    mol, loc = b.atom_mol, b.atom_loc
    PAA = Pb[mol, loc, loc]
    GAA = (np.einsum("ncd,nabcd->nab", PAA, G1) - 0.5 * np.einsum("ncd,nacbd->nab", PAA, G1))
    Fb = F.reshape(b.nmol, m, 4, m, 4).transpose(0, 1, 3, 2, 4)  # view
    pm, li, lj = b.pair_mol, b.atom_loc[b.idxi], b.atom_loc[b.idxj]
    PA = Pb[pm, li, li]
    PB = Pb[pm, lj, lj]
    PAB = Pb[pm, li, lj]
    JA = np.einsum("ncd,nabcd->nab", PB, T)      # onto atom i from density on j
    JB = np.einsum("nab,nabcd->ncd", PA, T)      # onto atom j from density on i
    K = -0.5 * np.einsum("nbd,nabcd->nac", PAB, T)  # onto block (i, j)
    np.add.at(Fb, (mol, loc, loc), GAA)
    np.add.at(Fb, (pm, li, li), JA)
    np.add.at(Fb, (pm, lj, lj), JB)
    np.add.at(Fb, (pm, li, lj), K)
    np.add.at(Fb, (pm, lj, li), np.transpose(K, (0, 2, 1)))
    for a in np.nonzero(b.Z == 1)[0]:
        o = 4 * loc[a]
        F[mol[a], o + 1:o + 4, :] = 0.0
        F[mol[a], :, o + 1:o + 4] = 0.0
    return F


# ------------------------------------------------------------ pack/unpack
def orbital_index(b, mol):
    """Padded row index of each real orbital of molecule mol (pack.py:8-17 order).

    Heavy atoms keep their 4 slots, hydrogens keep their s slot only.
    """
    idx = []
    for a in np.nonzero(b.atom_mol == mol)[0]:
        o = 4 * b.atom_loc[a]
        idx.extend([o, o + 1, o + 2, o + 3] if b.Z[a] > 1 else [o])
    return np.asarray(idx, dtype=np.int64)


# ----------------------------------------------------------- density step
def density_eigh(b, F):
    """sym_eig_trunc (seqm/seqm_functions/diag.py:110-241): e (padded with 0), P, C per molecule."""
    nmol, n = F.shape[0], F.shape[1]
    P = np.zeros_like(F)
    e = np.zeros((nmol, n))
    Cs = []
    for mo in range(nmol):
        idx = orbital_index(b, mo)
        f = F[mo][np.ix_(idx, idx)]
        w, V = np.linalg.eigh(f)
        k = int(b.nocc[mo])
        Pp = 2.0 * V[:, :k] @ V[:, :k].T
        P[mo][np.ix_(idx, idx)] = Pp
        e[mo, :len(idx)] = w
        Cs.append(V)
    return e, P, Cs


def sp2(a, nocc, eps=1.0e-4, factor=2.0):
    """Trace-correcting SP2 purification (seqm/seqm_functions/SP2.py:9-85), FP64 branch."""
    eps = min(max(eps, 1.0e-7), 1.0e-3)
    a = np.asarray(a, dtype=np.float64)
    N, D, _ = a.shape
    aii = np.einsum("nii->ni", a)
    ri = np.abs(a).sum(2) - np.abs(aii)
    h1 = (aii - ri).min(1)
    hN = (aii + ri).max(1)
    X = (np.eye(D)[None] * hN[:, None, None] - a) / (hN - h1)[:, None, None]
    nocc = np.asarray(nocc, dtype=np.float64)
    err0 = np.abs(np.einsum("nii->n", X) - nocc)
    err1 = err0.copy()
    active = np.ones(N, bool)
    iters = np.zeros(N, dtype=np.int64)
    while active.any():
        for mo in np.nonzero(active)[0]:
            x = X[mo]
            x2 = x @ x
            t2 = np.trace(x2)
            t1 = np.trace(x)
            X[mo] = x2 if abs(t2 - nocc[mo]) < abs(2.0 * t1 - t2 - nocc[mo]) else 2.0 * x - x2
            iters[mo] += 1
            err1[mo] = err0[mo]
            err0[mo] = abs(np.trace(X[mo]) - nocc[mo])
            if err0[mo] < eps and err1[mo] < eps:
                active[mo] = False
    return factor * X, iters


def density_sp2(b, F, eps):
    """pack -> SP2 -> unpack, scf_loop.py:66-78 (per-molecule packed size, no zero padding)."""
    P = np.zeros_like(F)
    for mo in range(F.shape[0]):
        idx = orbital_index(b, mo)
        f = F[mo][np.ix_(idx, idx)]
        x, _ = sp2(f[None], [b.nocc[mo]], eps)
        P[mo][np.ix_(idx, idx)] = x[0]
    return P


# ------------------------------------------------------------- energies
def elec_energy(P, F, H):
    """0.5 * sum P*(H + F), seqm/seqm_functions/energy.py:26-49 (H symmetric here)."""
    return 0.5 * np.einsum("nij,nij->n", P, H + F)


def elec_energy_xl(D, P, F, H):
    """XL-BOMD electronic energy sum D*F(P) - 0.5 sum (F(P) - H)*P, seqm/seqm_functions/energy.py:76-88
    (H symmetric here)."""
    return np.einsum("nij,nij->n", D, F) - 0.5 * np.einsum("nij,nij->n", F - H, P)


def pair_nuclear_energy(b, gam):
    """Core-core repulsion per pair, seqm/seqm_functions/energy.py:91-134 (MNDO/AM1/PM3)."""
    prm = b.prm
    rija = b.rij * A0
    alpha = prm["alpha"]
    t1 = TORE[b.ni] * TORE[b.nj] * gam
    XH = ((b.ni == 7) | (b.ni == 8)) & (b.nj == 1)
    tmp = np.exp(-alpha[b.ni] * rija)
    t2 = np.where(XH, tmp * rija, tmp)
    t3 = np.exp(-alpha[b.nj] * rija)
    E = t1 * (1.0 + t2 + t3)
    ng = N_GAUSS[b.method]
    if ng:
        t4 = TORE[b.ni] * TORE[b.nj] / rija
        t5 = np.zeros_like(rija)
        t6 = np.zeros_like(rija)
        for g in range(1, ng + 1):
            K, L, M = (prm[f"Gaussian{g}_{x}"] for x in "KLM")
            t5 += K[b.ni] * np.exp(-L[b.ni] * (rija - M[b.ni]) ** 2)
            t6 += K[b.nj] * np.exp(-L[b.nj] * (rija - M[b.nj]) ** 2)
        E = E + t4 * (t5 + t6)
    return E


def eiso(b):
    """Isolated-atom electronic energy, seqm/seqm_functions/energy.py:8-23."""
    prm, Z = b.prm, b.Z
    return (prm["U_ss"][Z] * USSC[Z] + prm["U_pp"][Z] * UPPC[Z] + prm["g_ss"][Z] * GSSC[Z]
            + prm["g_pp"][Z] * GPPC[Z] + prm["g_sp"][Z] * GSPC[Z] + prm["g_p2"][Z] * GP2C[Z]
            + prm["h_sp"][Z] * HSPC[Z])


# ------------------------------------------------------------ SCF drivers
MAX_ITER = 1000
DM_ERR_FACTOR, DM_ELEM_FACTOR, DIIS_FACTOR = 2.0, 15.0, 50.0  # scf_loop.py:50-52


def initial_density(b):
    """Diagonal guess: heavy tore/4 on s and p, H 1.0 on s (scf_loop.py:2066-2081)."""
    m = b.molsize
    P = np.zeros((b.nmol, 4 * m, 4 * m))
    for a in range(b.natoms):
        o = 4 * b.atom_loc[a]
        if b.Z[a] > 1:
            for k in range(4):
                P[b.atom_mol[a], o + k, o + k] = TORE[b.Z[a]] / 4.0
        else:
            P[b.atom_mol[a], o, o] = 1.0
    return P


class SCFState:
    pass


def _get_error(st, Pold, P, Enew, eps, diis_error=None):
    """4-way convergence test, scf_loop.py:106-147."""
    act = st.nc
    st.err[act] = Enew[act] - st.E[act]
    bad = np.abs(st.err) > eps
    if diis_error is not None:
        bad |= diis_error > DIIS_FACTOR * eps
    dm = act & ~bad
    for mo in np.nonzero(dm)[0]:
        dP = P[mo] - Pold[mo]
        st.dm_err[mo] = np.sqrt((dP * dP).sum()) / st.msize[mo]
        st.dm_elem[mo] = np.abs(dP).max()
    return bad | (st.dm_err > eps * DM_ERR_FACTOR) | (st.dm_elem > eps * DM_ELEM_FACTOR)


def _adaptive_mix(k, Pprev, Pcur, old2):
    """scf_loop.py:361-420 restated for a batch of active molecules."""
    third = k % 3 == 0
    damp = 0.05 if k > 4 else 1.0e10
    dprev = np.einsum("nii->ni", Pprev).copy()
    dcur = np.einsum("nii->ni", Pcur).copy()
    if third:
        d1 = dcur - dprev
        d2 = dcur - 2.0 * dprev + old2
        num = (d1 * d1).sum(1)
        den = (d2 * d2).sum(1)
        valid = (den > 0) & (num < 100.0 * den)
        fac = np.zeros_like(num)
        fac[valid] = np.sqrt(num[valid] / den[valid])
        Pmix = (1.0 + fac)[:, None, None] * Pcur - fac[:, None, None] * Pprev
    else:
        fac = np.zeros(Pcur.shape[0])
        Pmix = Pcur.copy()
    delta = dcur - dprev
    cap = np.abs(delta) > damp
    dnew = np.where(cap, dprev + np.sign(delta) * damp, dcur + fac[:, None] * delta)
    dnew = np.clip(dnew, 0.0, 2.0)
    sum0 = dcur.sum(1)
    di = dnew
    for _ in range(20):
        sum2 = di.sum(1)
        large = sum2 > 1.0e-3
        sum3 = np.zeros_like(sum2)
        sum3[large] = sum0[large] / sum2[large]
        if np.all((~large) | (np.abs(sum3 - 1.0) <= 1.0e-5)):
            break
        scaled = np.maximum(di * sum3[:, None], 0.0)
        full = scaled > 2.0
        di = np.where(full, 2.0, scaled)
        sum0 = sum0 - full.sum(1) * 2.0
    n = Pmix.shape[1]
    Pmix[:, np.arange(n), np.arange(n)] = di
    return Pmix, dprev


def scf(b, eps, converger=(2,), density="eigh", sp2_eps=1e-4, P=None, H=None, T=None,
        max_iter=MAX_ITER, verbose=False):
    """SCF on the padded batch: scf_forward0/1/2 (scf_loop.py:164-347, 424-635, 639-1132).

    converger = (0, alpha) linear, (1,) adaptive, (2,) Pulay DIIS with the reference's
    batch-global history counter and reset.  Returns P, notconverged, iteration count.
    """
    if H is None or T is None:
        H, T, _, _, _ = hcore(b)
    G1 = one_center_tensor(b.Z, b.prm)
    if P is None:
        P = initial_density(b)
    P = P.copy()

    def dens(Fa, mols):
        sub = _sub(b, mols)
        if density == "eigh":
            return density_eigh(sub, Fa)[1]
        return density_sp2(sub, Fa, sp2_eps)

    st = SCFState()
    nmol = b.nmol
    st.nc = np.ones(nmol, bool)
    st.err = np.ones(nmol)
    st.dm_err = np.ones(nmol)
    st.dm_elem = np.ones(nmol)
    st.msize = (4 * b.nHeavy + 4 * b.nHydro).astype(float)
    F = fock(b, P, H, T, G1)
    st.E = elec_energy(P, F, H)
    Pold = np.zeros_like(P)
    kind = converger[0]
    it = 0
    if kind in (0, 1):
        alpha = converger[1] if kind == 0 else None
        old2 = np.zeros(P.shape[:2])
        krange = range(0, max_iter + 1) if kind == 0 else range(1, max_iter + 1)
        for k in krange:
            it += 1
            act = np.nonzero(st.nc)[0]
            Pn = dens(F[act], act)
            if kind == 0:
                Pold[act] = P[act]
                P[act] = alpha * P[act] + (1.0 - alpha) * Pn
            else:
                Pmix, dprev = _adaptive_mix(k, P[act], Pn, old2[act])
                Pold[act] = P[act]
                P[act] = Pmix
                old2[act] = dprev
            F = fock(b, P, H, T, G1)
            En = st.E.copy()
            En[act] = elec_energy(P[act], F[act], H[act])
            st.nc = _get_error(st, Pold, P, En, eps)
            st.E[st.nc] = En[st.nc]
            if not st.nc.any():
                break
        return P, st.nc, it
    # Pulay DIIS (scf_loop.py:924-1131)
    nF = 10
    n = P.shape[1]
    tri = np.triu_indices(n)
    FPPF = np.zeros((nmol, nF, len(tri[0])))
    EMAT = np.tril(np.eye(nF + 1) - 1.0)[None].repeat(nmol, 0)
    FOCK = np.zeros((nmol, nF, n, n))
    counter, cF = -1, 0
    diis_err = np.full(nmol, np.finfo(float).max)
    for k in range(0, max_iter + 1):
        if not st.nc.any():
            break
        it += 1
        act = np.nonzero(st.nc)[0]
        cF = min(cF + 1, nF)
        counter = (counter + 1) % nF
        FOCK[act, counter] = F[act]
        C = F[act] @ P[act] - P[act] @ F[act]
        Cp = C[:, tri[0], tri[1]]
        FPPF[act, counter] = Cp
        diis_err[act] = np.abs(Cp).max(1)
        EMAT[act, counter, :cF] = np.einsum("nk,njk->nj", Cp, FPPF[act, :cF])
        reset = False
        if cF >= 2:
            EV_ = EMAT[act, :cF + 1, :cF + 1].copy()
            den = np.maximum(EV_[:, counter, counter], 1e-15)
            EV_[:, :cF, :cF] /= den[:, None, None]
            Lw, Q = np.linalg.eigh(EV_)  # lower triangle, as torch.linalg.eigh default
            absL = np.abs(Lw)
            cond = absL.max(1) / absL.min(1)
            reset = bool(np.any(cond > 1e7))
            inv = np.where(absL > 1e-13, 1.0 / np.where(absL > 1e-13, Lw, 1.0), 0.0)
            coeff = -np.einsum("bki,bi,bi->bk", Q[:, :cF, :], inv, Q[:, -1, :])
            F[act] = np.einsum("bk,bkij->bij", coeff, FOCK[act, :cF])
        Pn = dens(F[act], act)
        Pold[act] = P[act]
        P[act] = 0.5 * P[act] + 0.5 * Pn if cF < 2 else Pn
        F = fock(b, P, H, T, G1)
        En = st.E.copy()
        En[act] = elec_energy(P[act], F[act], H[act])
        st.nc = _get_error(st, Pold, P, En, eps, diis_err)
        st.E[st.nc] = En[st.nc]
        if reset:
            counter, cF = -1, 0
            FPPF[:] = 0.0
            EMAT = np.tril(np.eye(nF + 1) - 1.0)[None].repeat(nmol, 0)
            FOCK[:] = 0.0
    return P, st.nc, it


def _sub(b, mols):
    """Light view of a batch restricted to molecules `mols` (for per-molecule density steps)."""
    s = Batch()
    sel = np.isin(b.atom_mol, mols)
    remap = {int(mo): k for k, mo in enumerate(mols)}
    s.atom_mol = np.array([remap[int(x)] for x in b.atom_mol[sel]], dtype=np.int64)
    s.atom_loc = b.atom_loc[sel]
    s.Z = b.Z[sel]
    s.nocc = b.nocc[mols]
    return s


def single_point(species, coords, method, eps=1e-8, converger=(2,), density="eigh", sp2_eps=1e-4,
                 charges=None, max_iter=MAX_ITER, param_dir=None):
    """Energies of a batch: scf_loop (scf_loop.py:2034-2394) + Energy.forward (basics.py:813-892,
    1220-1239).  Returns a dict of numpy arrays."""
    b = parse(species, coords, method, charges, param_dir)
    H, T, di, gam, mp = hcore(b)
    P, nc, it = scf(b, eps, converger, density, sp2_eps, H=H, T=T, max_iter=max_iter)
    F = fock(b, P, H, T)
    e, _, _ = density_eigh(b, F)
    Eelec = elec_energy(P, F, H)
    EnucAB = pair_nuclear_energy(b, gam)
    Enuc = np.zeros(b.nmol)
    np.add.at(Enuc, b.pair_mol, EnucAB)
    Etot = Eelec + Enuc
    Eiso_sum = np.zeros(b.nmol)
    np.add.at(Eiso_sum, b.atom_mol, eiso(b))
    eheat = np.zeros(b.nmol)
    np.add.at(eheat, b.atom_mol, EHEAT_KCAL[b.Z] / EV_KCAL)
    Hf = Etot - Eiso_sum + eheat
    lumo = b.nocc
    e_gap = e[np.arange(b.nmol), lumo] - e[np.arange(b.nmol), lumo - 1]
    return dict(Hf=Hf, Etot=Etot, Eelec=Eelec, Enuc=Enuc, Eiso=Eiso_sum, e=e, e_gap=e_gap, P=P, F=F,
                H=H, notconverged=nc, iterations=it, batch=b)


# ------------------------------------------------------------- forces
def energy_fixed_density(species, coords, method, P):
    """E(R) = Eelec(P fixed) + Enuc: the Hellmann-Feynman energy whose gradient the reference's
    autograd returns when scf_backward = 0 (scf_loop.py:1996-2031, 2297-2320; basics.py:1315-1336)."""
    b = parse(species, coords, method)
    H, T, _, gam, _ = hcore(b)
    F = fock(b, P, H, T)
    Enuc = np.zeros(b.nmol)
    np.add.at(Enuc, b.pair_mol, pair_nuclear_energy(b, gam))
    return elec_energy(P, F, H) + Enuc


def energy_xl_fixed(species, coords, method, D, P):
    """E(R) = sum D F(P) - 0.5 sum (F(P) - H) P + Enuc at fixed D and P: the XL-BOMD energy of
    EnergyXL (seqm/dynamics/xlbomd.py:147-199, 430-447) whose gradient ForceXL's backward returns
    (xlbomd.py:536-551; D and P carry no gradient)."""
    b = parse(species, coords, method)
    H, T, _, gam, _ = hcore(b)
    F = fock(b, P, H, T)
    Enuc = np.zeros(b.nmol)
    np.add.at(Enuc, b.pair_mol, pair_nuclear_energy(b, gam))
    return elec_energy_xl(D, P, F, H) + Enuc


def forces_fd(species, coords, method, P, h=1.0e-4, D=None):
    """Force = -dE/dR (eV/Angstrom) by a 4th-order central difference of energy_fixed_density
    (D None) or of energy_xl_fixed (D given)."""
    coords = np.asarray(coords, dtype=np.float64)
    species = np.asarray(species)
    f = np.zeros_like(coords)
    for mo in range(coords.shape[0]):
        for a in range(coords.shape[1]):
            if species[mo, a] == 0:
                continue
            for x in range(3):
                vals = []
                for s in (-2, -1, 1, 2):
                    c = coords.copy()
                    c[mo, a, x] += s * h
                    if D is None:
                        e = energy_fixed_density(species[mo:mo + 1], c[mo:mo + 1], method, P[mo:mo + 1])
                    else:
                        e = energy_xl_fixed(species[mo:mo + 1], c[mo:mo + 1], method, D[mo:mo + 1],
                                            P[mo:mo + 1])
                    vals.append(e[0])
                g = (vals[0] - 8.0 * vals[1] + 8.0 * vals[2] - vals[3]) / (12.0 * h)
                f[mo, a, x] = -g
    return f


# --------------------------------------------------------------- XL-BOMD
# Niklasson et al. JCP 130, 214109 (2009) coefficients kappa, alpha, c_0..c_k
# (seqm/MolecularDynamics.py:1382-1390)
XL_COEFFS = {
    3: [1.69, 150e-3, -2.0, 3.0, 0.0, -1.0],
    4: [1.75, 57e-3, -3.0, 6.0, -2.0, -2.0, 1.0],
    5: [1.82, 18e-3, -6.0, 14.0, -8.0, -3.0, 4.0, -1.0],
    6: [1.84, 5.5e-3, -14.0, 36.0, -27.0, -2.0, 12.0, -6.0, 1.0],
    7: [1.86, 1.6e-3, -36.0, 99.0, -88.0, 11.0, 32.0, -25.0, 8.0, -1.0],
    8: [1.88, 0.44e-3, -99.0, 286.0, -286.0, 78.0, 78.0, -90.0, 42.0, -10.0, 1.0],
    9: [1.89, 0.12e-3, -286.0, 858.0, -936.0, 364.0, 168.0, -300.0, 184.0, -63.0, 12.0, -1.0],
}


def xl_propagate(D, P, Pt, k, step, c=0.95):
    """One density propagation of XL_BOMD (seqm/MolecularDynamics.py:1392-1427, 1458-1461):
    P(n+1) = 2 P(n) - P(n-1) + kappa (D' - P(n)) + alpha sum_i c_i P(n-i), D' = c D + (1-c) P(n),
    written over the ring Pt (m = k+1 slots; slot (m-1-cindx') holds the density of step
    n - ((cindx - cindx') mod m)).  Returns (P_new, Pt) with Pt updated in place."""
    kappa, alpha = XL_COEFFS[k][0], XL_COEFFS[k][1]
    cs = np.asarray(XL_COEFFS[k][2:]) * alpha
    m = k + 1
    cindx = step % m
    # P(n-i) sits in slot (m - cindx + i) mod m (the previous step wrote P(n) to slot m - cindx);
    # at the start every slot holds the seed density.
    hist = [Pt[(m - 1 - ((cindx - 1 - i) % m))] for i in range(m)]
    Dp = c * D + (1.0 - c) * P
    Pn = 2.0 * hist[0] - hist[1] + kappa * (Dp - hist[0])
    for i in range(m):
        Pn = Pn + cs[i] * hist[i]
    Pt[m - 1 - cindx] = Pn
    return Pn, Pt


# ------------------------------------------------ thermal density / KSA-XL-BOMD
KB_EV = 8.61739e-5          # Boltzmann constant, eV/K (seqm/dynamics/xlbomd.py:214)
CANON_DM_PRT_ITER = 10      # recursion steps of Canon_DM_PRT in EnergyXL (seqm/dynamics/xlbomd.py:55)
FERMI_MAX_ITER = 64         # fermi_q.py:12
FERMI_OCC_TOL = 1.0e-9      # fermi_q.py:14 (float64)
FERMI_ENTROPY_EPS = 1.0e-14
FERMI_TINY = 1.0e-30


def _sigmoid(x):
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(-x))


def fermi_q(b, F, T, kB=KB_EV):
    """Fermi_Q (seqm/seqm_functions/fermi_q.py:8-72) on every molecule's block of real orbitals:
    all eigenpairs (e ascending, Q columns), the chemical potential by Newton steps from the
    HOMO/LUMO midpoint with the reference's batch-wide stop (every |nocc - sum f| <= 1e-9, at most
    64 passes; sum f counts occupations per orbital, f = sigmoid(-(e - mu) / kB T)), the density
    D = 2 Q diag(f) Q^T (padded) and the entropy S = -kB sum [f ln f + (1-f) ln(1-f)] over
    1e-14 < f < 1 - 1e-14.  Returns D, S (nmol,), and per molecule lists e, Q, f, mu."""
    beta = 1.0 / (kB * T)
    es, Qs, idxs = [], [], []
    for mo in range(b.nmol):
        idx = orbital_index(b, mo)
        w, V = np.linalg.eigh(F[mo][np.ix_(idx, idx)])
        es.append(w)
        Qs.append(V)
        idxs.append(idx)
    N = b.nocc.astype(np.float64)
    mu = np.array([0.5 * (es[m][int(N[m]) - 1] + es[m][int(N[m])]) for m in range(b.nmol)])
    for _ in range(FERMI_MAX_ITER):
        fs = [_sigmoid(-beta * (es[m] - mu[m])) for m in range(b.nmol)]
        occ = np.array([f.sum() for f in fs])
        docc = np.array([max((beta * f * (1.0 - f)).sum(), FERMI_TINY) for f in fs])
        err = np.abs(N - occ)
        if (err <= FERMI_OCC_TOL).all():
            break
        mu = mu + (N - occ) / docc
    D = np.zeros_like(F)
    S = np.zeros(b.nmol)
    for m in range(b.nmol):
        f, V, idx = fs[m], Qs[m], idxs[m]
        D[m][np.ix_(idx, idx)] = 2.0 * (V * f) @ V.T
        ok = (f > FERMI_ENTROPY_EPS) & (1.0 - f > FERMI_ENTROPY_EPS)
        p = np.where(ok, f, 0.5)
        S[m] = (-kB * (p * np.log(p) + (1.0 - p) * np.log(1.0 - p)) * ok).sum()
    return D, S, es, Qs, fs, mu


def canon_dm_prt(FO1, T, Q, e, mu0, m=CANON_DM_PRT_ITER, kB=KB_EV):
    """Canon_DM_PRT (seqm/seqm_functions/canon_dm_prt.py:6-39) on one block of real orbitals: the
    first-order density response to the perturbation FO1 by the recursive Fermi-operator
    expansion in the eigenbasis of the unperturbed Fock matrix (e, Q; chemical potential mu0),
    m recursion steps, with the chemical-potential shift that keeps the trace.  Returns the block
    P1 = Q P1' Q^T (per orbital occupation, without the factor 2 of the density)."""
    beta = 1.0 / (kB * T)
    cnst = 2.0 ** (-2 - m) * beta
    X = Q.T @ FO1 @ Q
    p0 = 0.5 - cnst * (e - mu0)
    P1 = -cnst * X
    for _ in range(m):
        p02 = p0 * p0
        dX1 = p0[:, None] * P1 + P1 * p0[None, :]
        iD0 = 1.0 / (2.0 * (p02 - p0) + 1.0)
        p0 = iD0 * p02
        P1 = iD0[:, None] * (dX1 + 2.0 * (P1 - dX1) * p0[None, :])
    dpdmu = beta * p0 * (1.0 - p0)
    dmu1 = -np.trace(P1) / dpdmu.sum()
    P1 = P1 + np.diag(dpdmu) * dmu1
    return Q @ P1 @ Q.T


def g_matrix(b, V, T, G1=None):
    """G(V): the two-electron part of the Fock matrix of a density V, no core Hamiltonian
    (seqm/seqm_functions/G_XL_LR.py:7-908): fock with H = 0."""
    return fock(b, V, np.zeros_like(V), T, G1)


def ksa_rank_kernel(b, D, P, Qs, es, mus, T_el, Tt, rank, err_threshold, kB=KB_EV, K0=1.0):
    """The rank-m Krylov approximation of the XL-BOMD kernel in EnergyXL's max_rank branch
    (seqm/dynamics/xlbomd.py:243-339): Krylov vectors V_k (modified Gram-Schmidt in the Frobenius
    product, normalised), their responses W_k = K0 (P1(G(V_k)) - V_k) with P1 from canon_dm_prt,
    the least-squares fit of K0 (D - P) by the W_k (batch-wide loop while k < rank - 1 and the
    largest relative residual exceeds err_threshold).  Returns dP2dt2 = -sum_k alpha_k V_k (padded),
    the residuals and the rank used."""
    fro = lambda A, B: np.einsum("nij,nij->n", A, B)
    dDS = K0 * (D - P)
    dDS_norm = np.sqrt(fro(dDS, dDS))
    Vs, Ws = [], []
    dW = dDS
    k = -1
    Error = np.full(b.nmol, 10.0)
    idxs = [orbital_index(b, mo) for mo in range(b.nmol)]

    def fit(r):
        alpha = np.zeros((b.nmol, r))
        for mo in range(b.nmol):
            O = np.array([[np.sum(Ws[i][mo] * Ws[j][mo]) for j in range(r)] for i in range(r)])
            rhs = np.array([np.sum(Ws[i][mo] * dDS[mo]) for i in range(r)])
            alpha[mo] = np.linalg.solve(O, rhs)
        return alpha

    while k < rank - 1 and Error.max() > err_threshold:
        k += 1
        Vk = dW.copy()
        for j in range(k):
            Vk = Vk - fro(Vk, Vs[j])[:, None, None] * Vs[j]
        Vk = Vk / np.sqrt(fro(Vk, Vk))[:, None, None]
        Vs.append(Vk)
        FO1 = g_matrix(b, Vk, Tt)
        PO1 = np.zeros_like(Vk)
        for mo in range(b.nmol):
            ix = np.ix_(idxs[mo], idxs[mo])
            PO1[mo][ix] = canon_dm_prt(FO1[mo][ix], T_el, Qs[mo], es[mo], mus[mo], kB=kB)
        Wk = K0 * (PO1 - Vk)
        Ws.append(Wk)
        dW = Wk
        alpha = fit(k + 1)
        res = np.einsum("nr,rnij->nij", alpha, np.array(Ws)) - dDS
        Error = np.sqrt(fro(res, res)) / dDS_norm
    alpha = fit(k + 1)
    dP2dt2 = -np.einsum("nr,rnij->nij", alpha, np.array(Vs))
    return dP2dt2, Error, k + 1


def ksa_xl_propagate(dP2dt2, P, Pt, k, step):
    """KSA_XL_BOMD._propagate_P (seqm/MolecularDynamics.py:1615-1619): P(n+1) = kappa (dP2dt2 + P)
    + the same history terms as XL_BOMD: xl_propagate with D' = dP2dt2 + P and no mixing."""
    return xl_propagate(dP2dt2 + P, P, Pt, k, step, c=1.0)
