/*
 * plum_gpu.h -- C ABI of the MI355X trial-move energy engine.
 *
 * Drop-in boundary for the energy calls of nuwapi/Plum's ForceField
 * (src/force_field/force_field.h:207-294).  Plain C: opaque context, plain
 * pointers and sizes, status-code returns (0 = ok, < 0 = error, message via
 * plum_last_error), no exceptions across the boundary.  A context owns its
 * HIP streams: one for the host-path calls and, once the device loop or the
 * surrogate is used, a second for those; work queued on the first is ordered
 * before the device loop, and every call that returns results synchronizes.
 * Inputs stay caller-owned; device buffers are context-owned.
 *
 * Units and conventions are the reference's: kT = 1 energy unit, lengths in
 * sigma, `alpha` is the Ewald kappa^2 (sigma^-2), the Ewald cutoffs are
 * derived from alpha exactly as potential_ewald_coul.cc:43-75 does, the
 * energy difference is evaluated on trial coordinates and returned per term,
 * and the 1e8 sentinel of kVeryLargeEnergy marks a forbidden move.
 */
#ifndef PLUM_GPU_H_
#define PLUM_GPU_H_

#ifdef __cplusplus
extern "C" {
#endif

#define PLUM_MAX_TYPES 16

typedef struct plum_ctx plum_ctx;

enum plum_status {
  PLUM_OK = 0,
  PLUM_E_ARG = -1,          /* invalid argument or input */
  PLUM_E_HIP = -2,          /* HIP runtime error (no device, launch failure) */
  PLUM_E_NOMEM = -3,        /* device allocation failed */
  PLUM_E_STATE = -4,        /* call out of order (e.g. finalize without dE) */
  PLUM_E_CAPACITY = -5,     /* a fixed-capacity structure overflowed */
  PLUM_E_UNSUPPORTED = -6,  /* combination the engine does not implement */
  PLUM_E_PARSE = -7         /* malformed run.in / crd / top input */
};

/* Arithmetic of the pair sums. */
enum plum_precision {
  PLUM_FP64 = 0,       /* FP64 everywhere: parity path */
  PLUM_FP32_PAIR = 1   /* FP32 pair math (real space, LJ, k-space phases),
                          FP64 accumulation, FP64 S(k), FP64 acceptance */
};

/* Force-field parameters: the values run.in carries for the potentials used
   by ForceField::Initialize (force_field.cc:46-195).  Bead types are indices
   into the per-type tables; the caller maps bead symbols to indices. */
typedef struct {
  double box[3];        /* simulation box, unpadded (ReadTop) */
  int npbc;             /* 2 (slit, z walls) or 3 (bulk) */
  double beta;
  int use_pair, use_ewald, use_bond, use_ext;
  /* TruncatedLJ (potential_truncated_lj.cc:20-47) */
  double lj_cutoff;     /* < 0 selects the WCA form */
  int n_types;
  double lj_sigma[PLUM_MAX_TYPES], lj_eps[PLUM_MAX_TYPES];
  /* Ewald Coul (potential_ewald_coul.cc:33-36) */
  double lB, dielectric, alpha;
  int dipole_correction;
  /* Spring (potential_spring.cc:16) */
  double k_bond, r0;
  /* TruncatedLJWall (potential_truncated_lj_wall.cc:13-44) */
  double wall_cut, wall_sigma_w, wall_eps_w;
  double wall_sigma[PLUM_MAX_TYPES], wall_eps[PLUM_MAX_TYPES];
  int type_kind[PLUM_MAX_TYPES];   /* 0 plain, 1 grafted "L", 2 grafted "R" */
  /* engine options (no reference counterpart) */
  int precision;        /* enum plum_precision */
  int dipole_exact;     /* 0: reference dipole lag (default); 1: exact dE */
  int cell_div;         /* cells per search radius and dimension; 0 = auto */
} plum_params;

/* Per-term energies; layout shared with the CPU oracle's term order. */
typedef struct {
  double lj, real, kspace, self, dipole, bond, wall;
} plum_terms;

typedef struct {
  double box_e[3];      /* Ewald box (z padded with the dipole correction) */
  double vol_e;
  double real_cutoff;
  int real_cell[3];
  double repl_cutoff;
  int repl_cell[3];
  int n_k_full;         /* k vectors with 0 < k^2 <= repl_cutoff */
  int n_k_half;         /* stored half space (k and -k paired) */
  int n_cells[3];       /* real-space cell grid */
  int cell_cap;         /* slots per cell */
  int n_beads, n_mols, max_mol_len;
} plum_info;

/* ---- context ---------------------------------------------------------- */

/* Replaces ForceField::Initialize + InitializeEnergy
   (force_field.cc:46-363, :365-405) for the energy state: uploads the
   configuration (xyz interleaved x0,y0,z0,x1,..; molecule m owns beads
   mol_start[m] .. mol_start[m+1]-1), derives the Ewald cutoffs, builds the
   k-vector table, S(k), the cell list and all running totals on the GPU. */
int plum_create(const plum_params* params, int n_beads, const double* xyz,
                const double* q, const int* type, int n_mols,
                const int* mol_start, plum_ctx** out);
int plum_destroy(plum_ctx* ctx);

/* Derived Ewald cutoffs and engine sizes. */
int plum_get_info(const plum_ctx* ctx, plum_info* out);

/* From-scratch energy of the current configuration, per term, computed on
   the GPU.  Replaces the EnergyInitialization pass of each potential
   (potential_pair.cc:55-100, potential_ewald.cc:176-230,
   potential_bond.cc:16-27, potential_external.cc:48-61). */
int plum_energy_full(plum_ctx* ctx, plum_terms* out);

/* Replaces ForceField::EnergyDifference (force_field.cc:407-434): molecule
   `mol` moves; its beads idx[0..n_moved) get trial coordinates trial_xyz
   (interleaved).  Returns the per-term dE and the total with the
   reference's early exits (LJ >= 1e8, then wall >= 1e8, stop adding). */
int plum_energy_difference(plum_ctx* ctx, int mol, int n_moved, const int* idx,
                           const double* trial_xyz, plum_terms* dE,
                           double* total);

/* Replaces ForceField::FinalizeEnergies (force_field.cc:436-451) together
   with Simulation's commit/rollback of the moved molecule
   (simulation.cc:336-351). */
int plum_finalize(plum_ctx* ctx, int accept);

/* Running totals: ForceField::Tot{Pair,Ewald,Bond,Ext}Energy
   (force_field.cc:1201-1219), split per term. */
int plum_totals(plum_ctx* ctx, plum_terms* out);

/* Rebuilds S(k), the cell list and the dipole moment from scratch and
   resets the running totals to the from-scratch energy; `drift` (may be
   NULL) receives running - scratch before the reset.  No reference
   counterpart: the reference keeps incremental totals only. */
int plum_resync(plum_ctx* ctx, plum_terms* drift);

/* Current (committed) coordinates, interleaved, n_beads*3 doubles. */
int plum_get_positions(plum_ctx* ctx, double* xyz);

/* Message of the last failing call on this thread. */
const char* plum_last_error(void);

/* ---- GPU-resident Monte Carlo ----------------------------------------- */

/* Move-mix and run options of the device-resident trial-move loop.  The
   proposal generators restate Molecule::{BeadTranslate, COMTranslate, Pivot,
   Crankshaft, RandomReptation} (molecule.cc:103-312) and the selection of
   Simulation::TranslationalMove (simulation.cc:241-355), drawing from a
   counter-based Philox4x32-10 stream instead of mt19937. */
typedef struct {
  double move_prob[5];  /* bead, COM, pivot, crankshaft, reptation */
  double move_size;
  int phantom, grafted;
  double bond_len;      /* EqBondLen or rigid bond length */
  int vary_bond;        /* +-10% bond length in pivot / reptation */
  unsigned long long seed;
  int resync_every;     /* from-scratch resync period in moves (0 = never) */
} plum_mc_params;

typedef struct {
  long long attempted[5], accepted[5];
  long long moves;
  double seconds;       /* wall time of the run (host clock, synchronized) */
  plum_terms totals;    /* running totals at the end */
} plum_mc_stats;

/* Runs n_moves trial moves entirely on the GPU (proposal, dE, Metropolis,
   commit) on the context's stream, in speculative batches: up to `pmax`
   proposals on distinct molecules are evaluated against the same state in
   one launch and then accepted or rejected in order, each dE corrected by
   the exact cross terms of the proposals accepted before it (k space:
   2 sum_k c_k Re(dS_p* dS_q); LJ and real space: the pairs between the two
   moved sets), so the chain is the sequential one. */
int plum_mc_run(plum_ctx* ctx, const plum_mc_params* mc, long long n_moves,
                plum_mc_stats* stats);

/* One decided proposal of plum_mc_run: what Simulation::TranslationalMove's
   Metropolis test (simulation.cc:325-334) was given -- the per-term dE of
   ForceField::EnergyDifference (force_field.cc:407-434) with the exact cross
   terms of the proposals accepted before it in the batch, and the reference's
   early exits -- and its decision.  For replay checks of the device loop. */
typedef struct {
  unsigned long long move;   /* move number (Philox counter) */
  int mol, n_moved, move_type, accepted;
  int evaluated;             /* the exact dE was computed (0: a first-stage
                                rejection of delayed acceptance) */
  int early;                 /* an early exit (LJ or wall >= 1e8) */
  int first_stage_pass;      /* 1 without the surrogate */
  int lo, hi;                /* moved index range */
  plum_terms dE;
  double total;
  double dE_surrogate;       /* first-stage dE (0 without the surrogate) */
} plum_decision;

/* Arms a bounded record of the next `capacity` proposals plum_mc_run decides
   on this context (0 disarms); the count restarts at 0.  No reference
   counterpart. */
int plum_mc_record(plum_ctx* ctx, int capacity);
/* Copies up to `cap` recorded decisions in decision order; idx / xyz
   receive each one's moved bead indices and trial coordinates in slots of
   plum_max_mol_len(ctx) beads (cap * that many ints / 3 doubles; unused
   slots -1 / 0).  n_decided (may be NULL) receives the number of proposals
   decided since the record was armed, which may exceed what was kept.
   Returns the number of records copied or a negative status. */
int plum_mc_get_records(plum_ctx* ctx, int cap, plum_decision* out, int* idx, double* xyz,
                        long long* n_decided);

/* Average device duration (microseconds, HIP events on the context stream)
   of each engine kernel over the last plum_mc_run with profiling enabled.
   names/avg_us/count arrays of capacity `cap`; returns the number filled. */
int plum_set_profiling(plum_ctx* ctx, int on);
int plum_kernel_times(plum_ctx* ctx, int cap, const char** names,
                      double* avg_us, long long* count);

/* ---- batched replicas -------------------------------------------------- */

/* R independent systems (contexts created with plum_create, same precision)
   advanced together: every launch carries a speculative batch of every
   replica, and `batches_per_graph` batches are captured in a HIP graph and
   replayed.  No reference counterpart: the reference runs one system per
   process (simulation.cc:216-355 per replica).  The contexts must outlive the
   batch and must not be used on their own while a batch call runs. */
typedef struct plum_batch plum_batch;

typedef struct {
  int blocks_per_cu;     /* k_compute_spec work-groups per compute unit */
  int split_div;         /* candidate slots per share of a real-space query: one
                            block per share on the one-move path, one wave per share
                            in speculative batches (four when a batch holds fewer than
                            16 proposals; one wave in all when a query has fewer than
                            1.5 split_div slots) */
  int batches_per_graph; /* speculative batches captured in one graph (also
                            the cell compaction period) */
  int pmax;              /* proposals per speculative batch, 1..16 (1..32 screened with the surrogate) */
  double sweep_us;       /* measured time per move of every system (plum_batch_tune) */
  int cell_div;          /* cells per search radius and dimension of every system, 1..8
                            (0 keeps each system's grid); a change rebuilds the cell lists */
} plum_batch_tuning;

typedef struct {
  long long moves;       /* trial moves done, all replicas */
  long long sweeps;      /* moves per replica requested */
  long long attempted[5], accepted[5];
  long long overflow_replicas;  /* replicas finished alone after a cell overflow */
  double seconds;        /* wall time of the batched part (host clock, synchronized) */
} plum_batch_stats;

int plum_batch_create(plum_ctx* const* replicas, int n_replicas, plum_batch** out);
int plum_batch_destroy(plum_batch* batch);
/* mc[r] are replica r's move-mix parameters and seed; per (may be NULL)
   receives per-replica stats.  use_graph = 0 launches the same kernels
   directly (no graph). */
int plum_batch_mc_run(plum_batch* batch, const plum_mc_params* mc, long long moves_per_replica, int use_graph,
                      plum_batch_stats* total, plum_mc_stats* per);
/* Autotuner: for each candidate (cell grid, work-groups per CU, query split,
   proposals per batch) times `moves_per_trial` sweeps three times and keeps
   the candidate with the lowest median (the trial sweeps are ordinary
   moves). */
int plum_batch_tune(plum_batch* batch, const plum_mc_params* mc, int moves_per_trial, plum_batch_tuning* best);
int plum_batch_get_tuning(const plum_batch* batch, plum_batch_tuning* out);
int plum_batch_set_tuning(plum_batch* batch, const plum_batch_tuning* tuning);
/* Launch shape of plum_mc_run (one system). */
int plum_mc_get_tuning(plum_ctx* ctx, plum_batch_tuning* out);
int plum_mc_set_tuning(plum_ctx* ctx, const plum_batch_tuning* tuning);

/* ---- delayed acceptance with a learned surrogate ---------------------- */

/* No reference counterpart (the reference evaluates every trial move
   exactly).  With the surrogate on, plum_mc_run screens up to `screen`
   proposals per batch with a network estimate dE_s (exact spring and wall
   parts + a 64-128-128-128-1 tanh network over local radial-basis
   descriptors, BF16 matrix-core inference); the first stage accepts with
   min(1, exp(-beta dE_s)), and only its passes are evaluated exactly and
   accepted with min(1, exp(-beta (dE - dE_s))).  The sampled ensemble is the
   exact one; the network (FP32 master weights, Adam) is trained online on
   the exact dE of evaluated proposals. */
typedef struct {
  int screen;           /* proposals screened per batch, 1..32 */
  int fill;             /* exact slots (<= 8) also given to first-stage rejections, as training data */
  int train_every;      /* batches between online training steps; 0 = frozen network */
  int train_samples;    /* samples per training step (newest first) */
  int ring;             /* capacity of the training-sample ring */
  double lr;            /* Adam step size */
  double r_d;           /* descriptor radius (sigma) */
  double target_clip;   /* |target| clip of the training loss (kT) */
  unsigned long long seed;  /* weight initialisation */
} plum_surrogate_params;

typedef struct {
  long long s1_rejected;   /* first-stage rejections */
  long long s1_passed;     /* first-stage passes (exact evaluations decided) */
  long long train_evals;   /* first-stage rejections evaluated for training */
  long long stale_cuts;    /* batches ended early by a stale surrogate value */
  long long samples;       /* training samples collected */
  long long train_steps;
  double last_loss;        /* mean squared error of the last training step (kT^2) */
} plum_surrogate_stats;

#define PLUM_SURROGATE_PARAMS 41473   /* 64*128 + 128 + 2*(128*128 + 128) + 128 + 1 */

int plum_surrogate_enable(plum_ctx* ctx, const plum_surrogate_params* params);
int plum_surrogate_disable(plum_ctx* ctx);
int plum_surrogate_stats_get(plum_ctx* ctx, plum_surrogate_stats* out);
/* FP32 master weights: W1[64][128], b1, W2[128][128], b2, W3[128][128], b3,
   w4[128], b4 (PLUM_SURROGATE_PARAMS floats). */
int plum_surrogate_get_weights(plum_ctx* ctx, float* w, int n);
int plum_surrogate_set_weights(plum_ctx* ctx, const float* w, int n);
/* `steps` training steps on the newest samples; loss of the last one. */
int plum_surrogate_train(plum_ctx* ctx, int steps, double* loss);
/* Network outputs of `rows` descriptor rows (64 floats each): bf16 = 1 the
   matrix-core inference kernel, 0 the FP32 training forward pass. */
int plum_surrogate_forward(plum_ctx* ctx, const float* x, int rows, int bf16, float* f);
/* Loss and parameter gradient (FP32 kernels) of `samples` samples: sample i
   has rows [2 i maxm, 2 (i + 1) maxm) of x (its n[i] old beads first, its
   n[i] new beads from row maxm on), target[i]. */
int plum_surrogate_gradient(plum_ctx* ctx, const float* x, const double* target, const int* n, int samples,
                            double* loss, float* grad);
/* Descriptor rows (2 maxm x 64), old/new positions (2 maxm x 4) and moved
   count of proposal slot `slot` from the last batch (for checks). */
int plum_surrogate_debug_slot(plum_ctx* ctx, int slot, float* rows, double* pos, int* n);
int plum_max_mol_len(const plum_ctx* ctx);

/* ---- pressure --------------------------------------------------------- */

/* One volume perturbation of ForceField::CalcPressureVolScalingHSELSlit
   (pressure.cc:187-387) on the current configuration: the box is stretched
   along z by kDz = 1e-5 (every z scaled by 1 + kDz/Lz with the spring
   potential on, else each molecule shifted by kDz com_z/Lz; grafted "L"
   beads stay, "R" beads move by kDz) and dU is the energy change, over all
   pairs once (self pairs halved in the Ewald part; the first `phantom`
   molecules are single-bead surface sites, paired only from the first half
   of their list as the reference does; LJ between mobile beads), walls,
   springs and the slab dipole.  The caller accumulates the estimators
   (Yethiraj: <N/V> - <dU>/(A dz); de Miguel: ln <(1 + dz/Lz)^N e^{-beta dU}>
   / (beta A dz)). */
typedef struct {
  double dU;             /* total energy change of the stretch */
  double el, lj, wall, bond, dipole;   /* its parts (el: Ewald real + k space) */
  int n_mol;             /* molecules besides the surface sites */
  double vol;            /* Lx Ly Lz of the simulation box */
  double area_dz;        /* Lx Ly kDz */
} plum_vp;

int plum_vp_sample(plum_ctx* ctx, int phantom, plum_vp* out);

/* ---- configurational-bias chain growth (grand-canonical moves) --------- */

typedef struct {
  int n_trials;          /* K trial positions of this growth step */
  int grown;             /* beads of the trial chain placed so far */
  int exclude_lo, exclude_hi;  /* bead index range left out of the existing
                                  system (the chain being deleted and its
                                  counterions); exclude_lo > exclude_hi: none */
  int type_chain, type_ion;
  int charged;           /* each chain bead comes with a counterion */
  double q_chain, q_ion;
} plum_cbmc_request;

/* Replaces ForceField::BeadsEnergy (cbmc.cc:5-151) for all K trials of one
   growth step (the loop of CBMCFGenTrialBeads, cbmc.cc:161-212) in one
   pass on the GPU: trials = K chain-bead positions then, if charged, K
   counterion positions; grown = the chain beads placed so far then their
   counterions; energies[K] receives each trial's energy (1e8 when LJ or a
   wall forbids it), with the slab dipole term (DipoleEDiff,
   potential_ewald_coul.cc:473-530) when the context has it. */
int plum_cbmc_trials(plum_ctx* ctx, const plum_cbmc_request* req, const double* trials, const double* grown,
                     double* energies);

/* plum_cbmc_trials without any allocation.  No counterpart in the reference:
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops).  The request, the coordinates, the partial
   sums and the energies are staged in one 128 KiB file-scope device array per
   device instead of context buffers allocated on first use; limits: n_trials
   <= 64, grown <= 512, at most 32768 half-space k vectors (PLUM_E_CAPACITY
   beyond them).  Every call on a device shares that device's array: calls
   that overlap on one device race. */
int plum_cbmc_trials_noalloc(plum_ctx* ctx, const plum_cbmc_request* req, const double* trials,
                             const double* grown, double* energies);

typedef struct {
  int mode;              /* 0 insertion, 1 deletion, 2 Widom test insertion */
  int chain_len;         /* L beads (each with a counterion when charged) */
  int n_trials;          /* K trials per grown bead */
  int exclude_lo, exclude_hi;  /* deletion: the chain and its counterions */
  int type_chain, type_ion;
  int charged;
  double q_chain, q_ion;
} plum_growth_request;

/* One whole configurational-bias growth on the GPU with no host round trip
   between steps: the growth loops of CBMCFChainInsertion (cbmc.cc:222-266),
   CBMCFChainDeletion (cbmc.cc:361-400) and CalcChemicalPotentialF
   (cbmc.cc:472-524), with their trial energies (BeadsEnergy, cbmc.cc:5-151),
   weights, Rosenbluth product W and, for insertions, the choice of each bead.
   The caller supplies the draws in the reference's meaning: first[6] the
   first chain bead and its counterion (insertion, test insertion);
   disp/ions[(s K + t) 3] the bond vector and counterion position of trial t
   at step s = 1..L-1 (step 0's rows unused); sel[s] the uniform of the choice
   at step s; actual[2 L 3] the existing chain then its counterions
   (deletion: trial 0 of each step is the existing bead).  On return weight
   is W (the first bead's Boltzmann factor times prod_s W_s / K); stop_step is
   the step at which W reached 0 and the growth stopped (the reference returns
   there; -1 when it grew to the end); chain[2 L 3] (may be NULL) the grown
   chain then its counterions.  Calls that share a context serialize on its
   stream. */
int plum_cbmc_grow(plum_ctx* ctx, const plum_growth_request* req, const double* first, const double* disp,
                   const double* ions, const double* sel, const double* actual, double* weight, int* stop_step,
                   double* chain);

#ifdef __cplusplus
}
#endif
#endif
