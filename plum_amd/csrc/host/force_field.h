// Drop-in ForceField: the calling surface of the reference's ForceField
// (src/force_field/force_field.h:207-294) for the trial-move energy path,
// backed by the GPU engine through the C ABI of include/plum_gpu.h.
#pragma once
#include <iostream>
#include <map>
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

  RunIn r_;
  double box_[3] = {0, 0, 0};
  std::map<std::string, int> types_;
  std::vector<std::string> type_names_;
  std::vector<int> offset_;
  plum_ctx* ctx_ = nullptr;
  plum_terms last_dE_{};
  int n_chain_ = 0, n_cion_ = 0, n_aion_ = 0;
};

}  // namespace plum
