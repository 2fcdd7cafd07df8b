/*
 * plum_host.h -- C ABI of the host layer (libplum_host.so): the run.in
 * reader and a Monte Carlo driver with the reference's move semantics, on top
 * of the GPU engine of plum_gpu.h.  For callers that hand over a reference
 * input deck instead of building arrays themselves.
 */
#ifndef PLUM_HOST_H_
#define PLUM_HOST_H_

#include "plum_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plum_sim plum_sim;

/* What a run.in (plus its crd/top files) describes.  Replaces the input half
   of Simulation::Simulation (simulation.cc:23-201) and of
   ForceField::Initialize (force_field.cc:46-195); no GPU is touched. */
typedef struct {
  int n_beads, n_mols, n_types;
  int npbc, use_pair, use_ewald, use_bond, use_ext, use_gc, dipole_correction;
  double box[3];
  double beta, move_size, move_prob[5];
  long long steps;
  double lj_cutoff, lB, alpha, k_bond, r0, rigid_bond, wall_cut;
  int phantom, grafted;
} plum_runin_summary;

int plum_runin_parse(const char* run_in_path, plum_runin_summary* out);

/* Simulation built from a run.in path (crd/top resolved next to it), seeded
   explicitly (the reference seeds from the clock, simulation.cc:133-136). */
int plum_sim_create(const char* run_in_path, unsigned seed, plum_sim** out);
/* Simulation::Run without output files (simulation.cc:216-404): `steps`
   steps on the host RNG stream, each a translational trial move (energy
   difference on the GPU) or, with grand-canonical moves on, every gc_freq-th
   draw on average a chain insertion or deletion by configurational bias
   (cbmc.cc:213-441; each growth step's trial energies in one GPU pass). */
int plum_sim_run(plum_sim* sim, long long steps, long long attempted[5], long long accepted[5]);
/* The step sequence of plum_sim_run (the GC coin on the host stream), with
   the translational moves between two exchanges run as one batch by the
   engine's device loop, plum_mc_run, on its own counter-based streams: the same Markov
   chain, not the same draws.  No reference counterpart.  seconds[0]: wall
   time in the device loop, [1]: in exchanges (cumulative). */
int plum_sim_run_device(plum_sim* sim, long long steps, long long attempted[5], long long accepted[5],
                        double seconds[2]);
/* Grand-canonical counters: [0] insertions, [1] deletions; and the current
   number of beads and molecules (they change with accepted exchanges). */
int plum_sim_gc_stats(plum_sim* sim, long long attempted[2], long long accepted[2], int* n_beads, int* n_mols);
/* Two-stage grand-canonical exchange (delayed acceptance; no counterpart in
   the reference, whose exchange is mode 0).  Mode 2: an insertion draws its
   first chain bead and counterion as the reference's growth does, then
   accepts a first stage with min(1, r_1), r_1 = e^{beta mu} C w_1 theta
   (w_1: the first bead's exact Boltzmann factor, theta = e^{log_theta}: a
   fixed estimate of the rest of the Rosenbluth weight, C: the reference's
   insertion factor, cbmc.cc:268-300); only then grows the chain (the
   reference's draws and trial energies) and accepts with min(1, r / r_1),
   r = e^{beta mu} C W the reference's ratio.  A deletion's first stage is
   the inverse, C' / (e^{beta mu} w_1 theta) with the chain's own first bead.
   The chain sampled is the reference's ensemble; a first-stage rejection
   costs one trial energy instead of a growth.  Mode 1: the reference's
   exchange, recording ln(W / w_1) of every insertion growth, the statistic
   log_theta is fitted to.  Each stage draws its own uniform, so the stream
   is not the reference's in mode 2. */
int plum_sim_gc_prescreen(plum_sim* sim, int mode, double log_theta);
/* Counters of the two stages ([0] insertions, [1] deletions), the mean of
   ln(W / w_1) and the ln of the mean of W / w_1 over the recorded insertion
   growths (modes 1 and 2).  The acceptance is proportional to W, so the
   second, which the largest weights dominate, is the scale for theta. */
int plum_sim_gc_prescreen_stats(plum_sim* sim, long long s1_rejected[2], long long s1_passed[2],
                                double* mean_log_rest, double* log_mean_rest, long long* n_rest);
/* ForceField::CalcChemicalPotentialF (cbmc.cc:444-533): the mean weight of
   50 Widom test insertions of the exchanged chain species (or, without
   grand-canonical moves, of the first chain species, 30 trials per bead). */
int plum_sim_widom(plum_sim* sim, double* mean_weight);
/* Replaces the force sums of one call of ForceField::CalcPressureForceLJELSlit
   (pressure.cc:404-470) for the configuration given (phantom wall sites are
   molecules 0 .. phantom-1, the first half on the z = 0 wall): f[0..2] the
   LJ z forces on ions / chains / the other wall (PairForce,
   potential_truncated_lj.cc:87-121, plus half of BeadForceOnWall,
   potential_truncated_lj_wall.cc:136-214, on the system's beads), f[3..5]
   the Ewald real + reciprocal z forces likewise (PairForceZReal / Repl,
   potential_ewald_coul.cc:261-324), with the coefficients -1/2, +1/2 and -1
   of pressure.cc:425-429.  `ewald` is the engine's plum_info for the same
   parameters (k grid, cutoffs, images).  Host only; the caller does the
   area and call averaging of pressure.cc:472-487. */
int plum_slit_forces(const plum_params* p, const plum_info* ewald, int n_beads, const double* xyz,
                     const double* q, const int* type, int n_mols, const int* mol_start, int phantom,
                     double f[6]);
int plum_sim_energy(plum_sim* sim, plum_terms* running, plum_terms* from_scratch);
int plum_sim_positions(plum_sim* sim, double* xyz, int n_beads);
/* Simulation::RadiusOfGyration / EndToEndDistance (simulation.cc:406-501). */
int plum_sim_observables(plum_sim* sim, double* rg2, double* ree2);
/* Simulation::CalcRhoZ (simulation.cc:1029-1045) of the current
   configuration: out[3 b + s] for bins of width res along z, s = chain
   beads, cations, anions (beads per unit volume).  *n_bins receives the bin
   count; out may be NULL to ask for it. */
int plum_sim_density_profile(plum_sim* sim, double res, double* out, int max_bins, int* n_bins);
/* Simulation::PrintLastCrd / PrintLastTop (simulation.cc:935-1027): writes
   <prefix>_lastcrd.dat and <prefix>_lasttop.dat, from which a run resumes
   (a run.in naming them reads the configuration back). */
int plum_sim_write_last(plum_sim* sim, const char* prefix);
/* Simulation::Run with its sampling and output (simulation.cc:216-239;
   Sample :723-798, PrintStatHeader/PrintStat :685-867, PrintTraj :869-933,
   PrintLastCrd/PrintLastTop :935-1027, PrintLastRhoZ :1047-1064): steps
   steps writing <prefix>_stat.dat, <prefix>_traj_{p,c,a,n}.xyz,
   <prefix>_lastcrd.dat(2), <prefix>_lasttop.dat and <prefix>_lastrhoZ.dat(2)
   on the run.in's sample, stat and trajectory frequencies.  device != 0
   runs the translational moves between two exchanges or output steps as
   device batches.  The pressure column is nan, as the reference prints it
   without walls (pressure.cc:490-500); with walls its wall-force pressure is
   not built and the six columns are nan. */
int plum_sim_run_output(plum_sim* sim, long long steps, const char* prefix, int device);
/* ForceField::CalcPressureVolScalingHSELSlit (pressure.cc:187-387): one
   volume-perturbation sample of the current configuration (plum_vp_sample,
   the run.in's surface sites as phantom), added to the running estimators;
   sample (may be NULL) receives it, p_yethiraj / p_demiguel the averages so
   far in the reference's form (beta <N/V> - <dU>/(A dz), and
   beta/(A dz) ln <(1 + dz/Lz)^N e^{-beta dU}>), samples their count.  The
   reference computes these for boxes without walls and prints "nan"
   (pressure.cc:490-500); here they are returned. */
int plum_sim_pressure_sample(plum_sim* sim, plum_vp* sample, double* p_yethiraj, double* p_demiguel,
                             long long* samples);
int plum_sim_destroy(plum_sim* sim);
const char* plum_host_last_error(void);

/* Seeded synthetic system (chains as random walks of step r0 along Marsaglia
   directions, then counterions and salt pairs uniform; a bead landing within
   min_sep of another is redrawn).  No reference counterpart: the reference
   reads configurations from crd files (simulation.cc:582-646).  Types: 0 chain
   bead (2 for the grafted first bead when graft_z >= 0), 1 ion. */
typedef struct {
  int n_chains, chain_len, n_counterions, n_salt_pairs;
  double bead_q, counterion_q;
  double box[3];
  int npbc;
  double min_sep, r0;
  double graft_z;        /* < 0: free chains */
  double z_lo, z_hi;     /* growth / ion z range; z_hi <= z_lo: whole box */
  unsigned long long seed;
} plum_synth_spec;

int plum_synth_count(const plum_synth_spec* spec, int* n_beads, int* n_mols);
int plum_synth_system(const plum_synth_spec* spec, double* xyz, double* q, int* type, int* mol_start);

#ifdef __cplusplus
}
#endif
#endif
