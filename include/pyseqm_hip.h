/*
 * pyseqm_hip.h -- C ABI of the MI355X (gfx950) batched MNDO/AM1/PM3 closed-shell SCF kernels.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t (passed as void*),
 * enqueues its kernels on that stream and returns 0 on success or a negative PYSEQM_E* code.
 * No entry point synchronises the stream or allocates memory: every buffer, including
 * scratch, is owned by the caller, and entry points are reentrant (no global state), except
 * pyseqm_fock_noalloc, whose scratch is a library-owned device array (see its declaration).
 *
 * Batch layout ("packed", padding-free): molecule m owns the dense row-major norb_m x norb_m
 * block starting at mat_off[m] of every matrix buffer (P, F, H, ...).  Its atoms are stored
 * heavy atoms first (species sorted non-increasing, as the reference requires), a heavy atom
 * owns 4 consecutive orbitals (s, px, py, pz), a hydrogen owns 1 (s).  This is the orbital
 * order of the reference's packed matrices (seqm/seqm_functions/pack.py:8-17).
 *
 * Atom pairs i < j of one molecule are grouped in 8 x 8 "pair tiles" of three kinds:
 * heavy-heavy (XX), heavy-H (XH) and H-H (HH).  Each tile is processed by one 64-lane
 * wavefront, lane = 8 * r + c <-> pair (row0 + r, col0 + c).
 */
#ifndef PYSEQM_HIP_H
#define PYSEQM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PYSEQM_OK 0
#define PYSEQM_EINVAL (-1)      /* bad size / argument */
#define PYSEQM_ELAUNCH (-2)     /* kernel launch failed */
#define PYSEQM_EUNSUPPORTED (-3) /* element pair or method outside MNDO/AM1/PM3 sp scope */
#define PYSEQM_ETOOBIG (-4)     /* batch larger than a fixed-size library scratch buffer */

#define PYSEQM_NPAR 32 /* doubles per atom in the atom-parameter array (see below) */
#define PYSEQM_MAX_RED_BLOCKS 256 /* max blocks per molecule of the two-stage reductions */

/* Atom-parameter record, PYSEQM_NPAR doubles per atom.  Slots 0..23 are inputs (MOPAC
 * parameters gathered by Z, seqm/basics.py:406-448); 24..31 are filled by
 * pyseqm_atom_setup. */
enum {
  PAR_USS = 0, PAR_UPP, PAR_ZS, PAR_ZP, PAR_BS, PAR_BP, PAR_GSS, PAR_GSP, PAR_GPP, PAR_GP2,
  PAR_HSP, PAR_ALPHA, PAR_GAUSS = 12, /* K1 L1 M1 ... K4 L4 M4 */
  PAR_DD = 24, PAR_QQ, PAR_RHO0, PAR_RHO1, PAR_RHO2, PAR_TORE, PAR_QN, PAR_Z
};

enum { PYSEQM_MNDO = 0, PYSEQM_AM1 = 1, PYSEQM_PM3 = 2 };

/* Device-side description of one packed batch.  All pointers are device pointers. */
typedef struct {
  int nmol;
  int natoms;
  int ntile[3];              /* number of XX, XH, HH tiles */
  int npairs;                /* number of atom pairs (reference order) */
  const int* mol_atom_off;   /* [nmol+1] first atom of each molecule */
  const int* mol_norb;       /* [nmol] */
  const int* mol_nocc;       /* [nmol] */
  const long long* mol_mat_off; /* [nmol] offset (in doubles) of the molecule's matrix block */
  const int* mol_tile_off;   /* [3*(nmol+1)] tile range of each molecule, per tile kind */
  const int* atom_mol;       /* [natoms] */
  const int* atom_ao;        /* [natoms] first orbital of the atom inside its molecule */
  const int* atom_Z;         /* [natoms] */
  const int* tiles;          /* [4*ntiles_total]: mol, row0, col0, nrow | ncol<<8 | diag<<16 */
  const int* csr_off;        /* [natoms+1] per-atom list of tile partial slots */
  const int* csr_idx;        /* slot = tile * 16 + s (s < 8: row r = s, else column s - 8) */
  const long long* tile_woff; /* [ntiles_total] offset (doubles) of the tile's integrals in its kind's
                                buffer: NW rows (XX 100, XH 10, HH 1) of round_up(#valid pairs, 16)
                                doubles, valid lanes packed in lane order */
  int max_mol_atoms;         /* atoms of the largest molecule (0 = not given: the Fock build then
                                takes its tile-parallel kernels for every molecule) */
  int max_mol_slots;         /* max over molecules of 10 * heavy atoms + hydrogens */
  const int* mol_order;      /* [nmol] molecules by decreasing Fock work, or NULL (batch order) */
} pyseqm_layout_t;

/* Builds the pair tiles of a packed batch on the device: the pair enumeration of the reference
 * Parser (seqm/basics.py:219-403: pairs i < j of each molecule) grouped into 8 x 8 atom tiles of
 * the three kinds, as the tiles / tile_woff / csr_off / csr_idx fields of pyseqm_layout_t hold them.
 *   mol_atom_off [nmol+1], mol_nheavy [nmol]: atom offsets and heavy-atom counts (atoms of a
 *     molecule stored heavy first);
 *   mol_tile_off [3*(nmol+1)], ntile [3] (host array): each molecule's first tile per kind and the
 *     tile counts (closed forms the caller computes in O(nmol));
 *   outputs tiles [4*ntiles], tile_woff [ntiles], w_size [3] (doubles of each kind's integral
 *     buffer), csr_off [natoms+1], csr_idx [slots];
 *   scratch: ntiles + natoms/2 + 1 int64.
 * Records and slot lists come out in the host enumeration's order. */
int pyseqm_build_tiles(int nmol, int natoms, const int* mol_atom_off, const int* mol_nheavy, const int* mol_tile_off,
                       const int* ntile, int* tiles, long long* tile_woff, long long* w_size, int* csr_off,
                       int* csr_idx, long long* scratch, void* stream);

/* Derived per-atom quantities: dd, qq (charge separations), rho0/1/2 (additive terms),
 * tore, qn.  Replaces two_elec_two_center_int.py:117-147 + cal_par.py:11-280. */
int pyseqm_atom_setup(const pyseqm_layout_t* L, double* atom_par, void* stream);

/* Pair integrals for one geometry, fused: local-frame two-electron integrals, rotation to
 * the molecular frame (w), STO overlaps, core Hamiltonian blocks and the core-core
 * repulsion.  Replaces hcore() (seqm/seqm_functions/hcore.py:9-179) together with
 * two_elec_two_center_int (two_elec_two_center_int.py:98-283), diatom_overlap_matrix_PM6_SP
 * (diat_overlap_PM6_SP.py:6-444) and pair_nuclear_energy (energy.py:91-134).
 *   coords   [natoms*3] Angstrom
 *   w_xx     integral k*10+l of the pair on lane `lane` of tile t at
 *            w[tile_woff[t] + (k*10 + l)*S_t + rank(lane)], S_t = round_up(#valid lanes, 16),
 *            rank = position of the lane among the tile's valid lanes in lane order;
 *            size sum over XX tiles of 100*S_t doubles
 *   w_xh     same with 10 integrals per pair (sum 10*S_t)
 *   w_hh     same with 1 integral per pair (sum S_t)
 *   H        packed core Hamiltonian (full symmetric blocks)
 *   part     [ntiles_total*16*10] scratch: per-tile row/column partial sums
 *   enuc_tile[ntiles_total] per-tile core-core repulsion sums
 *   enuc_mol [nmol] output: per-molecule core-core repulsion (eV)
 *   enuc_pair[npairs] optional (may be NULL): per-pair repulsion in reference pair order,
 *   pair_ref [ntiles_total*64] maps tile lane -> reference pair index (or -1). */
int pyseqm_pair_integrals(const pyseqm_layout_t* L, int method, const double* coords,
                          const double* atom_par, double* w_xx, double* w_xh, double* w_hh,
                          double* H, double* part, double* enuc_tile, double* enuc_mol,
                          double* enuc_pair, const int* pair_ref, void* stream);

/* Forces -dE/dR (eV/Angstrom, [natoms*3]) at fixed density P: per atom pair, a 4th-order central
 * difference (step h Angstrom) of the pair's Hellmann-Feynman energy 2 P_ij.H_ij + P_ii.e1b +
 * P_jj.e2a + P_ii P_jj (ii|jj) - 1/2 P_ij P_ij (ij|ij) + E_nuc(ij) -- the energy the reference
 * differentiates by autograd with scf_backward = 0 (basics.py:1315-1336, scf_loop.py:1996-2031).
 * part: tile partial scratch as in pyseqm_pair_integrals. */
int pyseqm_pair_forces(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                       const double* P, double h, double* part, double* forces, void* stream);

/* Energy gradient with respect to the atom parameters at fixed density P: out[natoms][2][24] holds
 * dEelec/d(slot) and dEnuc/d(slot) for each atom's 24 input slots (PAR_USS .. the Gaussians).  At a
 * converged SCF density this is the exact gradient of Etot, the one the reference's autograd
 * returns for learned parameters (basics.py:781-789, Molecule.py:82-88) with scf_backward = 0
 * (scf_loop.py:1996-2031).  Pair terms: 4th-order central differences of each pair's energy in
 * z_s, z_p, beta_s, beta_p, g_ss, g_pp, g_p2, h_sp (through the multipoles, overlaps and
 * resonance integrals), alpha and the Gaussian terms, step hrel * max(|value|, 0.1), overlap
 * branches pinned; one-center terms (U_ss, U_pp, g_ss, g_sp, g_pp, g_p2, h_sp) in closed form.
 *   rec: pyseqm_param_grad_workspace(L) doubles; part: tile partial scratch as in
 *   pyseqm_pair_integrals. */
long long pyseqm_param_grad_workspace(const pyseqm_layout_t* L);
int pyseqm_param_grad(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                      const double* P, double hrel, double* rec, double* part, double* out, void* stream);
/* Vector-Jacobian product of the Fock matrix with the atom parameters: out[a][0][k] =
 * d(sum Y F(theta; P))/d(slot k of atom a) at fixed P and symmetric Y (out[a][1][.] = 0), the
 * contraction (dP_out/dtheta)^T u of the SCF implicit backward after the density response
 * (scf_loop.py:1779-1993, agrad(Pout, gv, grad_outputs=u) over M, w, gss .. hsp and on to the
 * parameters).  Same kernels, steps and workspace as pyseqm_param_grad. */
int pyseqm_fock_param_vjp(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                          const double* P, const double* Y, double hrel, double* rec, double* part, double* out,
                          void* stream);

/* Core-core repulsion E_nuc(i, j) of an explicit pair list, pair_nuclear_energy
 * (seqm/seqm_functions/energy.py:91-134): idxi/idxj [npairs] atom rows of atom_par (PAR_Z, PAR_TORE,
 * PAR_ALPHA and the Gaussian K/L/M slots are read), rij [npairs] in bohr, gam [npairs] = (ss|ss). */
int pyseqm_core_repulsion_pairs(long long npairs, int method, const int* idxi, const int* idxj, const double* rij,
                                const double* gam, const double* atom_par, double* enuc_pair, void* stream);

/* XL-BOMD forces: -d/dR of E(D, P) = sum D F(P) - 1/2 sum (F(P) - H) P + E_nuc at fixed D and P
 * (the energy of EnergyXL, energy.py:76-88, differentiated by ForceXL's backward,
 * seqm/dynamics/xlbomd.py:536-551).  D: density from F(P); P: propagated density field. */
int pyseqm_pair_forces_xl(const pyseqm_layout_t* L, int method, const double* coords, const double* atom_par,
                          const double* D, const double* P, double h, double* part, double* forces,
                          void* stream);

/* Fock build F = H + G(P) with the electronic energy 0.5*sum P*(H+F) per molecule fused in.
 * Replaces fock() (seqm/seqm_functions/fock.py:132-347) and elec_energy (energy.py:26-49).
 *   active [nmol] (may be NULL = all): molecules with active[m] == 0 are skipped.
 *   part [ntiles_total*16*10], etile [ntiles_total], eatom [natoms]: scratch (untouched when
 *     the layout's max_mol_atoms is 1 .. PYSEQM_FOCK_MOL_ATOMS: one workgroup per molecule then
 *     builds F and Eelec with the partial sums in LDS, sized by max_mol_slots; a molecule larger
 *     than those fields state is skipped with eelec[m] = NaN, nothing written out of bounds).
 *   eelec [nmol] output. */
#define PYSEQM_FOCK_MOL_ATOMS 24
int pyseqm_fock(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx,
                const double* w_xh, const double* w_hh, const double* P, const double* H,
                double* F, double* part, double* etile, double* eatom, double* eelec,
                const unsigned char* active, void* stream);

/* Fock build and Eelec that allocate nothing and take no scratch argument: the same kernels as
 * pyseqm_fock, with part / etile / eatom in a library-owned device array of
 * PYSEQM_FOCK_SCRATCH_DOUBLES doubles (48 MiB; one per GPU, held until the process ends, its
 * address looked up once per device on the first call with that device current).  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Returns
 * PYSEQM_ETOOBIG, before any HIP call, when pyseqm_fock_scratch_doubles(L) exceeds the array
 * (48 MiB holds a 3072-atom / 6144-orbital water cluster).  Calls share the array, so calls that
 * overlap on one GPU (concurrent streams) race; the scratch keeps the last call's contents and
 * needs no clearing (every slot a build reads it writes first). */
#define PYSEQM_FOCK_SCRATCH_DOUBLES (6LL * 1024 * 1024)
long long pyseqm_fock_scratch_doubles(const pyseqm_layout_t* L);
int pyseqm_fock_noalloc(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx,
                        const double* w_xh, const double* w_hh, const double* P, const double* H,
                        double* F, double* eelec, const unsigned char* active, void* stream);

/* ------------------------------------------------------------------ SCF loop plumbing
 * The functions below work on a list of active molecules idx[0..nact) (device int array): the
 * reference's `notconverged` masks (scf_loop.py:974-1131).  `red` is reduction scratch of at
 * least nmol * PYSEQM_MAX_RED_BLOCKS * 16 doubles; nb (1..PYSEQM_MAX_RED_BLOCKS) is the number of
 * blocks per molecule of the two-stage reductions. */

/* Density step, FP64 diagonalisation path (diag.py:110-241 sym_eig_trunc): copy each active
 * packed Fock matrix into a zero-padded nmax x nmax block whose padded diagonal lies above the
 * Gershgorin range (diag.py:172-192), for a batched symmetric eigensolver. */
int pyseqm_pad_for_eigh(const pyseqm_layout_t* L, const double* F, const int* idx, int nact, int nmax,
                        double* out, double* red, int nb, void* stream);
/* FP64 density step by diagonalisation with one workgroup per listed molecule, for blocks of up to
 * 192 orbitals: P = 2 C_occ C_occ^T of each packed Fock block F (sym_eig_trunc, diag.py:110-241,
 * and the density of make_Pnew, scf_loop.py:55-103), written to the packed P, with no padding.
 * Householder tridiagonalisation with the block in registers, multisection for the nocc lowest
 * eigenvalues, inverse iteration, Gram-Schmidt inside clusters of close eigenvalues,
 * back-transformation.  nmax / nocc_max must be at least every listed block's norb / nocc (they
 * size the LDS and pick the kernel variant; pyseqm_eigh_density_lds_ok(nmax, nocc_max) must hold); ws
 * holds pyseqm_eigh_density_workspace(nact, nmax) doubles; fail [1] counts vectors that
 * Gram-Schmidt reduced to zero (dependent vectors; it stays 0). */
long long pyseqm_eigh_density_workspace(int nact, int nmax);
int pyseqm_eigh_density_lds_ok(int norb, int nocc);
int pyseqm_eigh_density(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int nocc_max,
                        const double* F, double* P, double* ws, int* fail, void* stream);
/* P = 2 C_occ C_occ^T from the padded eigenvectors (columns) (diag.py:230-232).  V holds nact
 * blocks of nmax x nmax; nmax must be at least the orbital count of every listed molecule. */
int pyseqm_density_from_evecs(const pyseqm_layout_t* L, const double* V, int nmax, const int* idx,
                              int nact, double* P, void* stream);

/* Density mixing: Pold = P; P = alpha P + (1 - alpha) Pnew (scf_loop.py:266-267, 1052-1056). */
int pyseqm_mix(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, double alpha, double* P,
               const double* Pnew, double* Pold, void* stream);
/* Adaptive mixing, scf_converger [1] (adaptive_mix / compute_fac, scf_loop.py:350-420, as
 * scf_forward1 applies it at iteration k >= 1, scf_loop.py:516-548): Pold = P; off the diagonal
 * P = (1 + FAC) Pnew - FAC P every third iteration (else Pnew); the diagonal damped, clamped to
 * [0, 2] and renormalised to tr Pnew in the reference's batch-wide passes; old2 (the diagonal of P
 * two steps back) updated.  old2 holds one value per orbital, molecule m's from orb_off[m];
 * fac [nmol] and mask [1] are scratch.  nmax <= 8192.  No host synchronisation. */
int pyseqm_adaptive_mix(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int k, double* P,
                        const double* Pnew, double* Pold, double* old2, const long long* orb_off, double* fac,
                        unsigned* mask, void* stream);
/* dst = src on the upper triangle (c >= r) of the active blocks: the DIIS Fock ring keeps only
 * the upper triangles current (the matrices are symmetric; pyseqm_diis_combine reads only them). */
int pyseqm_copy_active(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* src,
                       double* dst, void* stream);

/* Convergence test get_error (scf_loop.py:106-147): |dE| <= eps, max|FP-PF| <= diis_tol,
 * ||dP||_F/(4 natoms) <= 2 eps, max|dP| <= 15 eps.  Writes notconv[m] (1 = not converged),
 * moves e_new into e_old for unconverged molecules, optional metrics[4*m..] = (dE, diis, dm, dm_max). */
int pyseqm_scf_check(const pyseqm_layout_t* L, const int* idx, int nact, const double* P,
                     const double* Pold, const double* e_new, double* e_old, const double* diis_err,
                     double eps, double diis_tol, unsigned char* notconv, double* metrics, double* red,
                     int nb, void* stream);

/* Ragged FP64 GEMM C = A B per active molecule (LDS-tiled FMA kernel, small matrices). */
int pyseqm_gemm(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* A,
                const double* B, double* C, void* stream);
/* FP64 MFMA GEMM C = A B with B symmetric.  off/lds: per-molecule offset and leading dimension
 * (lds NULL = packed, ld = n).  mode bit 0: McWeeny epilogue C = 2(3A - 2AB); bit 1: symmetric
 * product, only upper tiles are computed and mirrored. */
int pyseqm_dgemm_sym(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* off,
                     const int* lds, const double* A, const double* B, double* C, int mode, void* stream);

/* C = A B for the listed molecules' packed blocks (B symmetric) on bf16x3 MFMA with FP32
 * accumulation; FP64 in and out.  No reference counterpart of its own: it is the reduced-precision
 * product of the DIIS residual below for the mixed-precision density step. */
int pyseqm_gemm_bf16x3(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* A,
                       const double* B, double* C, void* stream);

/* Pulay DIIS (scf_loop.py:997-1035): residual C = FP - PF into a history slot and max|C|;
 * B-matrix row B[m][slot][j] = <C_slot, C_j> (packed upper triangle); F = sum_j c[m][j] F_j.
 * mixed != 0 forms F P on bf16x3 MFMA (mixed-precision density step), else on FP64 MFMA. */
int pyseqm_diis_residual(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* F,
                         const double* P, double* X, double* Cout, double* diis_err, double* red, int nb,
                         int mixed, void* stream);
/* The DIIS residual of the mixed-precision density step (scf_loop.py:997-1002): C = F P - P F on
 * bf16x3 MFMA over the upper-triangle tiles only (edge tb = 64 when nmax <= 256, else 128); writes
 * the strict upper triangle of
 * each listed block of Cout (packed layout; the lower triangle is left as it was) and max|C| to
 * diis_err[m].  Fx, Px: FP32 scratch in the SP2 X layout (xoff, xld), overwritten with zero-padded
 * copies of F and P; tmax: scratch of nact * tstride doubles, tstride >= nt (nt + 1) / 2 with
 * nt = ceil(nmax / tb). */
int pyseqm_diis_commutator_bf16x3(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* F,
                                  const double* P, const long long* xoff, const int* xld, float* Fx, float* Px,
                                  double* Cout, double* diis_err, double* tmax, int tstride, void* stream);
int pyseqm_diis_dots(const pyseqm_layout_t* L, const int* idx, int nact, const double* ring,
                     long long ring_stride, int slot, int count, int nfock, double* Bmat, double* red, int nb,
                     void* stream);
/* DIIS coefficients per listed molecule from the bordered B matrix Bmat[m][(nfock+1)^2] (lower
 * triangle; the count x count block normalised by B[slot][slot]), count + 1 <= 16: FP64 Jacobi
 * eigendecomposition and the eigenvalue pseudo-inverse of scf_loop.py:1015-1031; coef[m][0..count)
 * (stride nfock) and reset[m] = 1 when max|lambda| / min|lambda| > 1e7. */
int pyseqm_diis_solve(const pyseqm_layout_t* L, const int* idx, int nact, int nfock, int count, int slot,
                      const double* Bmat, double* coef, unsigned char* reset, void* stream);
/* F = sum_j coef[j] ring[j] (scf_loop.py:1033-1035), full symmetric F formed from the ring
 * entries' upper triangles. */
int pyseqm_diis_combine(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* ring,
                        long long ring_stride, int count, const double* coef, int cstride, double* F,
                        void* stream);

/* ------------------------------------------------------------------ SP2 density step
 * Trace-correcting SP2 (SP2.py:9-85) in mixed precision: X (FP32, per-molecule zero-padded
 * blocks at xoff[m] with leading dimension xld[m], a multiple of 32), X^2 on bf16x3 MFMA with FP32
 * accumulation, per-molecule state[8*m..] = (tr X, tr X^2, idempotency errors e = tr X - tr X^2
 * of this and the two previous iterates, next branch, done, iters) on the device.  A molecule
 * stops when its trace error is below eps and e_k >= 8 e_{k-2}^2 with e_{k-2} < 3/16 (the
 * quadratic decrease of exact arithmetic has ended), or at max_iter. */
int pyseqm_sp2_init(const pyseqm_layout_t* L, const double* F, const int* idx, int nact,
                    const long long* xoff, const int* xld, float* X, double* state, double* red, int nb,
                    void* stream);
/* One SP2 iteration for every active, not yet stopped molecule: Xout = Xin^2 or 2 Xin - Xin^2,
 * then the device-side branch / stop decision; *ndone counts molecules that stopped.
 * The product runs over upper-triangle output tiles of edge tb = 64 when nmax <= 256, else 128;
 * part holds 2 doubles per tile and molecule: part_stride >= nt (nt + 1), nt = ceil(nmax / tb)
 * (returns PYSEQM_EINVAL otherwise). */
int pyseqm_sp2_iterate(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* xoff,
                       const int* xld, const float* Xin, float* Xout, double* state, double* part,
                       int part_stride, double eps, int max_iter, int* ndone, void* stream);
/* `count` iterations it0 .. it0+count-1 in one call; iteration it reads Xa when it is even (else Xb)
 * and writes the other buffer. */
int pyseqm_sp2_iterate_many(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* xoff,
                            const int* xld, float* Xa, float* Xb, double* state, double* part, int part_stride,
                            double eps, int max_iter, int* ndone, int it0, int count, void* stream);
/* McWeeny purification of the final X (SP2.py:9-85 followed by the FP64 correction this path
 * adds), P = 2 (3 X^2 - 2 X^3) written to the packed density P (offsets mol_mat_off), computed as
 * Y2 = X X on the FP64 MFMA GEMM and P = 2 (Y2 - 2 X D), D = Y2 - X, with X D on bf16x3 MFMA (D is
 * of the size of X's idempotency error, so P keeps FP64 accuracy: ~1e-11 for a converged X).
 * X is each molecule's FP32 block in Xa or Xb by the parity of its iteration count, widened
 * exactly to FP64 as it is loaded; D (FP32) overwrites the molecule's block in the other buffer;
 * Y2 is an FP64 scratch with the X layout (upper-triangle tiles written). */
int pyseqm_sp2_mcweeny(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* xoff,
                       const int* xld, const double* state, float* Xa, float* Xb, double* Y2, double* P,
                       void* stream);

/* ------------------------------------------------------------------ thermal density / KSA-XL-BOMD
 * Every eigenpair of the listed blocks (the eigen-decomposition of Fermi_Q, fermi_q.py:27-30, via
 * sym_eig_trunc(eig_only=True), diag.py:110-241): Q in the packed layout (column j = eigenvector
 * j, ascending eigenvalues), evals[orb_off[m] + j].  The workgroup eigensolver with all n vectors:
 * nmax <= 192 and pyseqm_eigh_density_lds_ok(nmax, nmax); ws as for pyseqm_eigh_density. */
int pyseqm_eigh_vectors(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const double* F,
                        const long long* orb_off, double* evals, double* Q, double* ws, int* fail, void* stream);
/* Fermi_Q (fermi_q.py:8-72) after the eigen-decomposition: the chemical potential by Newton steps
 * from the HOMO/LUMO midpoint, stopping at the first step at which every listed molecule has
 * |nocc - sum f| <= 1e-9 (at most 64), occupations occ[orb_off[m] + j] = 1 / (1 + exp((e - mu) / kB T)),
 * mu[m], D = 2 Q diag(occ) Q^T (packed) and the entropy S[m] = -kB sum [f ln f + (1-f) ln(1-f)]
 * over 1e-14 < f < 1 - 1e-14.  Requires 1 <= nocc < norb <= 256 per molecule; work holds
 * pyseqm_fermi_workspace(nact) bytes. */
long long pyseqm_fermi_workspace(int nact);
int pyseqm_fermi_density(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* orb_off,
                         const double* evals, const double* Q, double T, double kB, double* occ, double* mu,
                         double* D, double* S, void* work, void* stream);
/* C = op(A) op(B) on every listed molecule's packed n x n block (ta / tb: transpose), FP64: the
 * eigenbasis transforms Q^T X Q and Q X Q^T of Canon_DM_PRT (canon_dm_prt.py:17, 37). */
int pyseqm_block_gemm(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int ta, int tb,
                      const double* A, const double* B, double* C, void* stream);
/* First-order closed-shell density response in the eigenbasis, in place on Y = Q^T X Q:
 * Y_ia = Y_ai <- 2 Y_ia / (e_i - e_a) for i < nocc <= a, zero elsewhere (the eigh backward inside
 * the SCF implicit backward, scf_loop.py:1779-1993).  Self-adjoint. */
int pyseqm_density_response_elem(const pyseqm_layout_t* L, const int* idx, int nact, int nmax,
                                 const long long* orb_off, const double* evals, double* Y, void* stream);
/* The canonical density-matrix perturbation recursion of Canon_DM_PRT (canon_dm_prt.py:18-35) in
 * place on Y = Q^T FO1 Q (eigenbasis; eigenvalues evals, chemical potential mu[m], `steps`
 * recursion steps), then the trace-keeping chemical-potential shift. */
int pyseqm_canon_prt(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, const long long* orb_off,
                     const double* evals, const double* mu, double T, double kB, int steps, double* Y, void* stream);
/* The Krylov fit of EnergyXL's max_rank branch (xlbomd.py:260-339): out[m] = sum of A * B over the
 * block (Frobenius product); Y <- Y + s c_m X (mode 0), Y / sqrt(c_m) (1), s (Y - X) (2), s c_m X (3),
 * Y + s X (4) with c_m = c[m * cstride]; alpha = O^{-1} rhs per molecule (O at m * R^2, rhs and
 * alpha at m * R, R = pyseqm_ksa_max_rank(), r <= R; partial pivoting). */
int pyseqm_block_dot(const pyseqm_layout_t* L, const int* idx, int nact, const double* A, const double* B,
                     double* out, void* stream);
int pyseqm_block_axpy(const pyseqm_layout_t* L, const int* idx, int nact, int nmax, int mode, double s,
                      const double* c, int cstride, const double* X, double* Y, void* stream);
int pyseqm_ksa_max_rank(void);
int pyseqm_small_solve(const int* idx, int nact, int r, const double* O, const double* rhs, double* alpha,
                       void* stream);

/* ------------------------------------------------------------------ XL-BOMD / MD
 * Density propagation of XL_BOMD._propagate_P (seqm/MolecularDynamics.py:1418-1427) plus the
 * ring update Pt[m-1-cindx] = P (1461) over `total` packed elements:
 *   P = coef_D (c D + (1 - c) P) + sum_{j<m} coef[j] Pt[j*pt_stride + .];  Pt[slot] = P.
 * coef: HOST array of m <= 11 doubles (the reference's coeff[cindx : cindx + m]). */
int pyseqm_xl_propagate(long long total, const double* D, double* P, double* Pt, long long pt_stride, int m,
                        const double* coef, double coef_D, double c, int slot, void* stream);
/* XL electronic energy per molecule, elec_energy_xl (energy.py:76-88):
 * eelec[m] = sum D F - 1/2 sum (F - H) P over the molecule's packed block (H full symmetric).
 * red: [nmol*nb] scratch, 1 <= nb <= PYSEQM_MAX_RED_BLOCKS. */
int pyseqm_xl_energy(const pyseqm_layout_t* L, const double* D, const double* P, const double* F,
                     const double* H, double* eelec, double* red, int nb, void* stream);
/* Velocity-Verlet half steps on packed (natoms, 3) arrays (MolecularDynamics.py:1449-1450,
 * 1503-1505): phase 0: v += dt/2 a, x += dt v;  phase 1: a = f minv acc_scale, v += dt/2 a.
 * minv [natoms]: inverse masses. */
int pyseqm_md_verlet(int natoms, int phase, double dt, double acc_scale, double* x, double* v, double* a,
                     const double* f, const double* minv, void* stream);

#ifdef __cplusplus
}
#endif
#endif
