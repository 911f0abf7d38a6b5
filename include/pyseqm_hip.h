/*
 * pyseqm_hip.h -- C ABI of the MI355X (gfx950) batched MNDO/AM1/PM3 closed-shell SCF kernels.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t (passed as void*),
 * enqueues its kernels on that stream and returns 0 on success or a negative PYSEQM_E* code.
 * No entry point synchronises the stream or allocates memory: every buffer, including
 * scratch, is owned by the caller.  Entry points are reentrant (no global state).
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

#define PYSEQM_NPAR 32 /* doubles per atom in the atom-parameter array (see below) */

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
} pyseqm_layout_t;

/* Derived per-atom quantities: dd, qq (charge separations), rho0/1/2 (additive terms),
 * tore, qn.  Replaces two_elec_two_center_int.py:117-147 + cal_par.py:11-280. */
int pyseqm_atom_setup(const pyseqm_layout_t* L, double* atom_par, void* stream);

/* Pair integrals for one geometry, fused: local-frame two-electron integrals, rotation to
 * the molecular frame (w), STO overlaps, core Hamiltonian blocks and the core-core
 * repulsion.  Replaces hcore() (seqm/seqm_functions/hcore.py:9-179) together with
 * two_elec_two_center_int (two_elec_two_center_int.py:98-283), diatom_overlap_matrix_PM6_SP
 * (diat_overlap_PM6_SP.py:6-444) and pair_nuclear_energy (energy.py:91-134).
 *   coords   [natoms*3] Angstrom
 *   w_xx     [ntile[0]*6400]  tile-major: w[(tile*100 + 10*k + l)*64 + lane]
 *   w_xh     [ntile[1]*640]
 *   w_hh     [ntile[2]*64]
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

/* Fock build F = H + G(P) with the electronic energy 0.5*sum P*(H+F) per molecule fused in.
 * Replaces fock() (seqm/seqm_functions/fock.py:132-347) and elec_energy (energy.py:26-49).
 *   active [nmol] (may be NULL = all): molecules with active[m] == 0 are skipped.
 *   part [ntiles_total*16*10], etile [ntiles_total], eatom [natoms]: scratch.
 *   eelec [nmol] output. */
int pyseqm_fock(const pyseqm_layout_t* L, const double* atom_par, const double* w_xx,
                const double* w_xh, const double* w_hh, const double* P, const double* H,
                double* F, double* part, double* etile, double* eatom, double* eelec,
                const unsigned char* active, void* stream);

#ifdef __cplusplus
}
#endif
#endif
