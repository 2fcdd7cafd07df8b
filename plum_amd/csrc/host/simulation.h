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
  // configuration as <prefix>_lastcrd.dat<suffix> and <prefix>_lasttop.dat,
  // each molecule shifted by whole boxes so that its centre lies in the box.
  void PrintLastCrd(const std::string& prefix, const std::string& suffix = "") const;
  void PrintLastTop(const std::string& prefix) const;
  // Simulation::RadiusOfGyrationXYZ (simulation.cc:437-481): mean squared
  // radius of gyration of the chains per axis.
  void RadiusOfGyrationXYZ(double out[3]) const;
  // The reference's Run loop with its sampling and output
  // (simulation.cc:216-239): after every step Sample, PrintStat, PrintTraj,
  // PrintLastCrd, PrintLastTop and PrintLastRhoZ, each on its own schedule,
  // into <prefix>_stat.dat, <prefix>_traj_{p,c,a,n}.xyz,
  // <prefix>_lastcrd.dat(2), <prefix>_lasttop.dat and
  // <prefix>_lastrhoZ.dat(2).  Steps are numbered on from the steps already
  // done.  With device = true the translational moves between two GC moves
  // or sampling points run as device batches (as RunDevice does).
  void RunWithOutput(long long steps, const std::string& prefix, bool device);
  // Running averages of Sample (simulation.cc:723-798).
  struct Averages {
    long long ff_count = 0, mol_count = 0;
    double pair = 0, ewald = 0, bond = 0, ext = 0, density = 0, chem_pot = 0;
    double rg = 0, rg_xyz[3] = {0, 0, 0}, ete = 0;
    std::vector<double> rho_z;   // 3 per bin: chain beads, cations (and neutral), anions
    double rho_z_res = 0.02;
    long long gc_attempted[2] = {0, 0}, gc_accepted[2] = {0, 0};   // after equilibration
  } avg;

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
  void device_moves(long long n);
  void sample(long long step);
  void print_stat_header(std::ostream& o) const;
  void print_stat(std::ostream& o, long long step) const;
  void print_traj(std::ostream* out[4], long long step) const;
  void print_last_rho_z(const std::string& prefix) const;
  int n_mol = 0, n_chain = 0, n_cion = 0, n_aion = 0, n_nion = 0;
};

}  // namespace plum
