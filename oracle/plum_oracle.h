/*
 * plum_oracle.h -- CPU FP64 restatement of the reference trial-move energy
 * path (nuwapi/Plum), used ONLY as the test oracle.
 *
 * This is test infrastructure: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product (plum_amd/) never links or calls it.
 *
 * Every function restates a reference routine; the citation is on the
 * definition in plum_oracle.c.  Pinning (what checks each function besides
 * itself) is listed in the header comment of plum_oracle.c.
 */
#ifndef PLUM_ORACLE_H_
#define PLUM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_TYPES 16

/* Bead kinds for the wall potential (symbol "L", "R" or anything else). */
enum { ORC_KIND_PLAIN = 0, ORC_KIND_L = 1, ORC_KIND_R = 2 };

/* Per-term energy record.  Order is fixed and shared with the product ABI. */
enum {
  ORC_T_LJ = 0, ORC_T_REAL = 1, ORC_T_KSPACE = 2, ORC_T_SELF = 3,
  ORC_T_DIPOLE = 4, ORC_T_BOND = 5, ORC_T_WALL = 6, ORC_NTERMS = 7
};

typedef struct {
  double box[3];      /* simulation box (unpadded) */
  int npbc;           /* 2 (slit) or 3 (bulk) */
  double beta;
  int use_pair, use_ewald, use_bond, use_ext;
  /* TruncatedLJ */
  double lj_cutoff;   /* < 0 : WCA */
  int n_types;
  double lj_sigma[ORC_MAX_TYPES], lj_eps[ORC_MAX_TYPES];
  /* Ewald Coul */
  double lB, dielectric, alpha;
  int dipole;
  /* Spring */
  double k_bond, r0;
  /* TruncatedLJWall */
  double wall_cut, wall_sigma_w, wall_eps_w;
  double wall_sigma[ORC_MAX_TYPES], wall_eps[ORC_MAX_TYPES];
  int type_kind[ORC_MAX_TYPES];
} orc_params;

typedef struct {
  double box_e[3];    /* Ewald box (z padded when dipole correction is on) */
  double vol_e;
  double real_cutoff;
  int real_cell[3];
  double repl_cutoff;
  int repl_cell[3];
  int n_k_full;       /* grid points with 0 < k2 <= repl_cutoff */
} orc_ewald_info;

typedef struct orc_sys orc_sys;

/* Creation copies the arrays.  xyz is interleaved (x0,y0,z0,x1,...).
   mol_start has n_mols+1 entries (beads of molecule m are
   mol_start[m] .. mol_start[m+1]-1, in file order). */
orc_sys* orc_create(const orc_params* p, int n_beads, const double* xyz,
                    const double* q, const int* type, int n_mols,
                    const int* mol_start);
/* As orc_create; init_energy = 0 skips the from-scratch energy pass (and so
   leaves the running totals at zero) for large systems that are only
   sampled. */
orc_sys* orc_create_ex(const orc_params* p, int n_beads, const double* xyz,
                       const double* q, const int* type, int n_mols,
                       const int* mol_start, int init_energy);
void orc_destroy(orc_sys* s);
void orc_get_ewald_info(const orc_sys* s, orc_ewald_info* out);

/* Standalone cutoff derivation (potential_ewald_coul.cc:43-75). */
void orc_derive_cutoffs(const double box_in[3], double lB, double alpha,
                        int dipole, orc_ewald_info* out);

/* Full (from-scratch) energy of the current configuration, per term. */
void orc_energy_full(orc_sys* s, double terms[ORC_NTERMS]);
/* k-space total in the structure-factor form (same sum, regrouped). */
double orc_kspace_total_sk(orc_sys* s);
/* Reference pair functions on the CURRENT coordinates of beads i, j. */
double orc_pair_lj(orc_sys* s, int i, int j);
double orc_pair_real(orc_sys* s, int i, int j);
double orc_pair_repl(orc_sys* s, int i, int j);
double orc_wall_bead(orc_sys* s, int i);

/* Trial move set-up: moved beads of molecule `mol` get trial coordinates. */
void orc_set_trial(orc_sys* s, int mol, int n_moved, const int* idx,
                   const double* trial_xyz);
/* ForceField::EnergyDifference with its early exits; per-term dE into
   terms[], total returned. */
double orc_energy_difference(orc_sys* s, int mol, double terms[ORC_NTERMS]);
/* ForceField::FinalizeEnergies + Simulation commit/rollback + UnsetMoved. */
void orc_finalize(orc_sys* s, int mol, int accept);
/* Running totals (incremental, like Tot*Energy). */
void orc_running_totals(const orc_sys* s, double terms[ORC_NTERMS]);
void orc_get_positions(const orc_sys* s, double* xyz);
int orc_n_beads(const orc_sys* s);

/* ---- std::mt19937 and the reference move generators ---------------- */
typedef struct { uint32_t mt[624]; int mti; } orc_mt19937;
void orc_mt_seed(orc_mt19937* r, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* r);
void orc_rand_sphere(orc_mt19937* r, double vec[3]);

/* Move generators: write trial coordinates of molecule `mol` into the
   system (all beads of the molecule; moved flags set as the reference does).
   Return the number of moved beads. */
int orc_move_bead_translate(orc_sys* s, int mol, double move_size, orc_mt19937* r);
int orc_move_com(orc_sys* s, int mol, double move_size, orc_mt19937* r);
int orc_move_pivot(orc_sys* s, int mol, double move_size, orc_mt19937* r,
                   double eq_bond_len, int vary_bond);
int orc_move_crankshaft(orc_sys* s, int mol, double move_size, orc_mt19937* r);
int orc_move_reptation(orc_sys* s, int mol, orc_mt19937* r, double eq_bond_len,
                       int vary_bond);
/* Export the moved list / trial coordinates of the pending move. */
int orc_get_moved(const orc_sys* s, int mol, int* idx, double* trial_xyz);

typedef struct {
  int attempted, move_type, mol, accepted, n_moved;
  double dE;
  double terms[ORC_NTERMS];
} orc_step_info;

/* One Simulation::Run() step without GC: the per-step draw, then
   Simulation::TranslationalMove. */
void orc_mc_step(orc_sys* s, orc_mt19937* r, const double move_prob[5],
                 double move_size, int phantom, int grafted, double bond_len,
                 int vary_bond, orc_step_info* info);
/* Simulation::TranslationalMove (simulation.cc:241-355) alone. */
void orc_translational_move(orc_sys* s, orc_mt19937* r, const double move_prob[5],
                            double move_size, int phantom, int grafted, double bond_len,
                            int vary_bond, orc_step_info* info);

/* Pairwise energy (LJ + real + k space, reference pair functions) of a bead
   of charge q and type t at position xyz against beads [0, n_partners):
   the inner loop of PotentialEwald::EnergyDifference for one moved bead
   (potential_ewald.cc:479-514) and of PotentialPair::EnergyDifference
   (potential_pair.cc:266-282).  Used to time the reference algorithm on a
   bounded sample. */
double orc_bead_energy_sample(orc_sys* s, const double xyz[3], double q, int t,
                              int n_partners);

/* Observables (simulation.cc:406-501). */
double orc_radius_of_gyration(const orc_sys* s);
double orc_end_to_end(const orc_sys* s);

/* ForceField::BeadsEnergy (cbmc.cc:5-151) for one trial (chain bead b1 and,
   if charged, its counterion b2); molecules ex_lo..ex_hi left out (none if
   ex_lo > ex_hi).  Pinned by the GPU trial energies only through this
   restatement of the reference loops (parity unpinned beyond the pair
   functions it calls, which are pinned). */
double orc_cbmc_dipole_ediff(const orc_sys* s, const double b1[3], double q1, const double b2[3], double q2,
                             int charged, const double* grown_c, const double* grown_i, int cur, int ex_lo,
                             int ex_hi);
double orc_cbmc_beads_energy(orc_sys* s, const double b1[3], int t1, double q1, const double b2[3],
                             int t2, double q2, int charged, const double* grown_c,
                             const double* grown_i, int cur, int ex_lo, int ex_hi);

/* Grand-canonical chain exchange with configurational bias.
   The exchanged chain has chain_len beads of charge bead_charge and type
   type_c; a charged chain comes with chain_len counterions of charge
   -bead_charge and the same type (force_field.cc:385-400).  bond: 1 harmonic
   spring (bond lengths from PotentialSpring::RandomBondLen), 2 rigid
   (rigid_bond), 0 none (length 0). */
typedef struct {
  double chem_pot, debroglie, bead_charge, rigid_bond;
  int chain_len, trials, type_c, bond;
  int phantom, coion, grafted, grafted_counterion;
} orc_gc;
/* state of gasdev's pair cache (misc.cc:111-133 keeps it in statics) */
typedef struct { int iset; double gset; } orc_gauss;

/* ForceField::CBMCFChainInsertion (cbmc.cc:213-335).  Returns 1 when the
   insertion is accepted; chain_out (2 chain_len x 3) receives the grown
   chain then its counterions; *weight the Rosenbluth weight. */
int orc_gc_insertion(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double* chain_out,
                     double* weight);
/* ForceField::CBMCFChainDeletion (cbmc.cc:337-441).  Returns the molecule
   index of the deleted chain (its counterions follow it) or -1. */
int orc_gc_insertion2(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double log_theta,
                      double* chain_out, double* weight_out, int* stage);
int orc_gc_deletion2(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double log_theta,
                     double* weight_out, int* stage);
int orc_gc_deletion(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double* weight);
/* ForceField::CalcChemicalPotentialF (cbmc.cc:444-533) with n_ins test
   insertions (mu_tot_ins, 50 in the reference). */
double orc_gc_widom(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, int n_ins);
double orc_gasdev(double mean, double stdev, orc_mt19937* r, orc_gauss* gs);
/* PotentialSpring::RandomBondLen (potential_spring.cc:50-61) with the
   system's k_bond, r0 and beta. */
double orc_random_bond_len(orc_sys* s, orc_mt19937* r, orc_gauss* gs);
/* One volume perturbation of ForceField::CalcPressureVolScalingHSELSlit
   (pressure.cc:187-387): out[0] dU, [1..5] its Ewald-pair, LJ-pair, wall,
   spring and dipole parts, [6] n_mol - phantom, [7] Lx Ly Lz, [8] Lx Ly kDz. */
void orc_vp_sample(orc_sys* s, int phantom, double out[9]);
/* The force sums of one call of ForceField::CalcPressureForceLJELSlit
   (pressure.cc:404-470) on the current coordinates, phantom wall sites first
   (half on each wall): f[0..2] LJ on ions / chains / the other wall, f[3..5]
   the Ewald real + reciprocal z forces likewise (before the area and call
   averaging of pressure.cc:472-487). */
void orc_slit_forces(orc_sys* s, int phantom, double f[6]);
/* PotentialTruncatedLJ::PairForce, PotentialEwaldCoul::PairForceZReal /
   PairForceZRepl (wall site a first) and PotentialTruncatedLJWall::
   BeadForceOnWall on the current coordinates of beads a, b. */
double orc_lj_force(orc_sys* s, int a, int b);
double orc_real_force_z(orc_sys* s, int a, int b);
double orc_repl_force_z(orc_sys* s, int a, int b);
double orc_wall_force(orc_sys* s, int a);

#ifdef __cplusplus
}
#endif
#endif
