// Drop-in ForceField: the calling surface of the reference's ForceField
// (src/force_field/force_field.h:207-294) for the trial-move energy path,
// backed by the GPU engine through the C ABI of include/plum_gpu.h.
#pragma once
#include <iostream>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "model.h"
#include "plum_gpu.h"
#include "runin.h"

namespace plum {

class ForceField {
 public:
  ForceField() = default;
  ~ForceField();
  ForceField(const ForceField&) = delete;
  ForceField& operator=(const ForceField&) = delete;

  // force_field.cc:46: reads the force-field part of run.in from `in`
  // (std::cin in the reference), assigns bead types and builds the GPU state.
  // Throws std::runtime_error on malformed input or engine failure.
  void Initialize(double beta, int npbc, double box_l[3], std::vector<Molecule>& mols, int phantom, int coion,
                  int grafted, int grafted_counterion, std::istream& in = std::cin);
  // Same, with the run.in values already parsed.
  void Initialize(const RunIn& r, double box_l[3], std::vector<Molecule>& mols);
  // force_field.cc:365: (re)builds the energy state from the molecules'
  // current coordinates.
  void InitializeEnergy(std::vector<Molecule>& mols);

  // force_field.cc:407: dE of the trial coordinates of the moved beads of
  // mols[moved_mol].
  double EnergyDifference(std::vector<Molecule>& mols, int moved_mol);
  // force_field.cc:436
  void FinalizeEnergies(std::vector<Molecule>& mols, bool accept, int moved_mol);

  // force_field.cc:1201-1219 (running totals)
  double TotPairEnergy();
  double TotEwaldEnergy();
  double TotBondEnergy();
  double TotExtEnergy();
  plum_terms LastDifference() const { return last_dE_; }
  plum_terms Totals();
  plum_terms FullEnergy();

  // Grand-canonical chain exchange with configurational bias (cbmc.cc);
  // the trial energies of each growth step are one plum_cbmc_trials call.
  // cbmc.cc:161: K trials around end_bead; returns their summed weight.
  double CBMCFGenTrialBeads(const Bead& end_bead, std::vector<Molecule>& mols, int current_len,
                            std::mt19937& rand_gen, int delete_id);
  // cbmc.cc:213: on acceptance the chain and its counterions are appended to
  // mols; the caller then calls EnergyInitForAddedMolecule.
  bool CBMCFChainInsertion(std::vector<Molecule>& mols, std::mt19937& rand_gen);
  // cbmc.cc:337: returns the index of the chain to delete (its counterions
  // follow it) or -1; the caller erases the molecules and then calls
  // UpdateMolCounts, which rebuilds the energy state.
  int CBMCFChainDeletion(std::vector<Molecule>& mols, std::mt19937& rand_gen);
  // cbmc.cc:444: Widom test insertions of the exchanged (or, without GC
  // moves, the first) chain species.
  double CalcChemicalPotentialF(std::vector<Molecule>& mols, std::mt19937& rand_gen);
  // Two-stage exchange (no counterpart in the reference; mode 2): the first
  // stage accepts on the first bead's exact Boltzmann factor w_1 times a fixed
  // estimate theta = e^{log_theta} of the rest of the growth's Rosenbluth
  // weight, r_1 = e^{beta mu} C w_1 theta (deletion: its inverse), so a
  // rejected exchange costs one trial energy instead of a growth; the second
  // stage takes the exact weight, min(1, r / r_1).  Mode 1 keeps the
  // reference's exchange and only records ln(W / w_1) of every insertion
  // growth (the statistic theta is fitted to); mode 0 is the reference's.
  void SetGCPrescreen(int mode, double log_theta);
  struct GCPrescreenStats {
    long long s1_rejected[2] = {0, 0}, s1_passed[2] = {0, 0};
    double sum_log_rest = 0;   // sum of ln(W / w_1) over recorded insertion growths
    double lse_rest = 0;       // ln sum of W / w_1 over them (running log-sum-exp)
    long long n_rest = 0;
    void record(double log_rest);
  };
  const GCPrescreenStats& PrescreenStats() const { return pre_; }
  // force_field.cc:1088: rebuilds the energy state with the appended molecules.
  void EnergyInitForAddedMolecule(std::vector<Molecule>& mols);
  const std::vector<double>& TrialWeights() const { return wts_; }
  // n translational moves on the device loop (plum_mc_run); the molecules'
  // coordinates are refreshed from the device afterwards.
  plum_mc_stats DeviceMoves(std::vector<Molecule>& mols, const plum_mc_params& mc, long long n);

  // pressure.cc:187-387: one volume perturbation of the current
  // configuration (plum_vp_sample on the GPU), accumulated as the reference
  // accumulates it; returns the sample.  The estimators in the reference's
  // form: Yethiraj (beta <N/V> - <dU> / (A dz)) and de Miguel
  // (beta / (A dz) ln <(1 + dz/Lz)^N e^{-beta dU}>).
  plum_vp CalcPressureVolScalingHSELSlit(std::vector<Molecule>& mols);
  double PressureYethiraj() const;
  double PressureDeMiguel() const;
  // ForceField::CalcPressureForceLJELSlit (pressure.cc:404-487): one sample
  // of the wall-force slit pressure from the current molecules (host,
  // plum_slit_forces), accumulated as the reference accumulates p_tensor.
  void CalcPressureForceLJELSlit(std::vector<Molecule>& mols);
  // ForceField::GetPressure (pressure.cc:490-500).
  std::string GetPressure() const;
  long long PressureSamples() const { return vp_z_; }

  bool UseGC() const { return r_.use_gc; }
  bool UsePairPot() const { return r_.use_pair; }
  bool UseEwaldPot() const { return r_.use_ewald; }
  bool UseBondPot() const { return r_.use_bond; }
  bool UseBondRigid() const { return r_.use_rigid_bond; }
  bool UseAnglePot() const { return r_.use_angle; }
  bool UseDihedPot() const { return r_.use_dihed; }
  bool UseExtPot() const { return r_.use_ext; }
  double RigidBondLen() const { return r_.rigid_bond; }
  double EqBondLen() const;
  int GCFrequency() const { return r_.gc_freq; }
  void UpdateMolCounts(std::vector<Molecule>& mols);

  plum_ctx* ctx() { return ctx_; }
  const RunIn& params() const { return r_; }
  int TypeOf(const std::string& symbol) const;
  int BeadIndex(int mol, int k) const { return offset_[mol] + k; }

 private:
  void build_params(plum_params& p, const double box_l[3]);
  void check(int rc, const char* what);
  void assign_types(std::vector<Molecule>& mols);
  void sync(std::vector<Molecule>& mols);
  double random_bond_len(std::mt19937& g);
  double gasdev(double mean, double stdev, std::mt19937& g);
  void trial_energies(int K, int cur, int delete_id, const double* tri, double* e);
  void draw_growth(std::mt19937& g, int mode, int last, bool sel_last, bool draw_first = true);
  double grow_device(std::vector<Molecule>& mols, std::mt19937& g, int mode, int delete_id,
                     bool draw_first = true);
  double insertion_factor(std::vector<Molecule>& mols);   // C of cbmc.cc:268-300
  double deletion_factor(std::vector<Molecule>& mols);    // C of cbmc.cc:403-441
  double first_bead_weight(const double first[6], int delete_id);
  bool insertion_two_stage(std::vector<Molecule>& mols, std::mt19937& g);
  int deletion_two_stage(std::vector<Molecule>& mols, std::mt19937& g, int del);
  void append_chain(std::vector<Molecule>& mols);
  void gc_counts(std::vector<Molecule>& mols, int& spc1, int& spc2);

  RunIn r_;
  double box_[3] = {0, 0, 0};
  std::map<std::string, int> types_;
  std::vector<std::string> type_names_;
  std::vector<int> offset_;
  plum_ctx* ctx_ = nullptr;
  plum_terms last_dE_{};
  int n_chain_ = 0, n_cion_ = 0, n_aion_ = 0;
  int n_beads_ = 0;        // beads of mols the energy state was built from
  bool ghost_ = false;     // an empty system carries one inert bead
  // exchanged species: chain_len beads of charge q_gc_, counterions of
  // charge -q_gc_, both of type gc_type_ (force_field.cc:385-400)
  int gc_len_ = 0, gc_trials_ = 0, gc_type_ = 0;
  double q_gc_ = 0;
  int gauss_iset_ = 0;     // gasdev's cached deviate (misc.cc:112-113)
  double gauss_gset_ = 0;
  std::vector<double> chain_, tri_, wts_, en_;
  std::vector<double> first_, disp_, ion_, sel_, act_;  // draws of one growth
  int gc_pre_mode_ = 0;      // SetGCPrescreen
  double gc_log_theta_ = 0;
  GCPrescreenStats pre_;
  // pressure accumulators (pressure.cc:187-387): samples, sum of N / V, sum
  // of dU, sum of (1 + dz/Lz)^N e^{-beta dU}, and the last A dz
  long long vp_z_ = 0;
  double vp_ideal_ = 0, vp_du_ = 0, vp_dm_ = 0, vp_adz_ = 0;
  long long slit_z_ = 0;        // calls of CalcPressureForceLJELSlit (the reference's vp_z there)
  double p_tensor_[12] = {0};   // pressure.cc:389-403
};

}  // namespace plum
