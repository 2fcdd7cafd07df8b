// Host Monte Carlo driver with the reference's move semantics
// (Simulation::Run / TranslationalMove / GCMove, simulation.cc:216-404, and the move
// generators of molecule.cc:103-312) on std::mt19937, calling the GPU
// engine through the drop-in ForceField.  Unlike the reference, the seed is
// a parameter (the reference seeds from the clock, simulation.cc:133-136).
#pragma once
#include <istream>
#include <random>
#include <string>
#include <vector>

#include "force_field.h"
#include "model.h"
#include "runin.h"

namespace plum {

// Move generators: write trial coordinates and moved flags of one molecule.
void BeadTranslate(Molecule& m, double move_size, std::mt19937& rng);
void COMTranslate(Molecule& m, double move_size, std::mt19937& rng);
void Pivot(Molecule& m, double move_size, std::mt19937& rng, double eq_bond_len, bool vary_bond);
void Crankshaft(Molecule& m, double move_size, std::mt19937& rng);
void RandomReptation(Molecule& m, std::mt19937& rng, double eq_bond_len, bool vary_bond);

class Simulation {
 public:
  // Reads the whole run.in from `in`; crd/top paths are taken relative to
  // base_dir when they are not absolute.  Throws on malformed input.
  Simulation(std::istream& in, unsigned seed, const std::string& base_dir = "");
  void Run(long long steps);
  void TranslationalMove();
  // simulation.cc:357-404: insertion or deletion of one exchanged chain
  void GCMove();
  // The same step sequence as Run (the GC coin on the host stream), with the
  // translational moves between two exchanges run as one batch by the
  // engine's device loop (plum_mc_run, its own counter-based streams): the
  // same Markov chain, not the same draws.
  void RunDevice(long long steps);
  double seconds_device = 0, seconds_gc = 0;
  double RadiusOfGyration() const;
  double EndToEndDistance() const;
  // Simulation::CalcRhoZ (simulation.cc:1029-1045) of the current
  // configuration: out[3 b + s], s = chain beads, cations, anions, in beads
  // per unit volume of the bins of width res; returns the bin count.
  int DensityProfile(double res, std::vector<double>& out) const;
  // Simulation::PrintLastCrd / PrintLastTop (simulation.cc:935-1027): the
  // configuration as <prefix>_lastcrd.dat and <prefix>_lasttop.dat, each
  // molecule shifted by whole boxes so that its centre lies in the box.
  void PrintLastCrd(const std::string& prefix) const;
  void PrintLastTop(const std::string& prefix) const;

  RunIn r;
  double box[3] = {0, 0, 0};
  std::vector<Molecule> mols;
  ForceField ff;
  std::mt19937 rng;
  long long attempted[5] = {0, 0, 0, 0, 0};
  long long accepted[5] = {0, 0, 0, 0, 0};
  long long steps_done = 0;
  long long gc_attempted[2] = {0, 0};  // insertions, deletions
  long long gc_accepted[2] = {0, 0};

 private:
  void update_counts();
  int n_mol = 0, n_chain = 0, n_cion = 0, n_aion = 0, n_nion = 0;
};

}  // namespace plum
