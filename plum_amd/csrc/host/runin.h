// Positional run.in reader (the reference's input format: `name value`
// pairs where only the order matters).  Grammar: Simulation ctor
// (simulation.cc:30-50), ForceField::Initialize (force_field.cc:65-195) and
// each potential's ReadParameters; topology (simulation.cc:610-646) and
// coordinate (simulation.cc:582-608) files.
#pragma once
#include <istream>
#include <string>
#include <vector>

#include "model.h"

namespace plum {

struct TypeParam {
  std::string symbol;
  double sigma = 0, eps = 0;
};

struct RunIn {
  // simulation block
  std::string crd_file, top_file, out_prefix;
  long long steps = 0, steps_eq = 0;
  long long sample_freq = 1, stat_out_freq = 1, traj_out_freq = 1;
  int npbc = 3;
  double beta = 1, move_size = 1;
  int calc_chem_pot = 0, phantom = 0, coion = 0, grafted = 0, grafted_counterion = 0;
  double move_prob[5] = {0, 0, 0, 0, 0};
  // force-field block
  double vp_el_res = 0, vp_bead_size = 0;
  int use_pair = 0, use_ewald = 0, use_bond = 0, use_rigid_bond = 0, use_angle = 0, use_dihed = 0;
  int use_ext = 0, use_gc = 0;
  double mu = 0, debroglie = 1;
  int gc_freq = 0, gc_chain_len = 0, gc_bead_charge = 0, cbmc_trials = 0;
  std::string gc_bead_symbol;
  // potentials
  std::string pair_name, ewald_name, bond_name, ext_name;
  double lj_cutoff = 0;
  std::vector<TypeParam> lj_types;
  double lB = 0, dielectric = 0, alpha = 0;
  int dipole_correction = 0;
  double k_bond = 0, r0 = 0;
  double rigid_bond = -1;
  double wall_cut = 0, wall_sigma = 0, wall_eps = 0;
  std::vector<TypeParam> wall_types;
};

// Each returns false and sets `err` on malformed input (the reference prints
// a message and calls exit(1) instead).
bool read_sim_block(std::istream& in, RunIn& r, std::string& err);
bool read_ff_block(std::istream& in, RunIn& r, std::string& err);
bool read_top(const std::string& path, int& n_bead, int& n_mol, double box[3],
              std::vector<Molecule>& mols, std::string& err);
bool read_crd(const std::string& path, std::vector<Molecule>& mols, int& id_counter, std::string& err);

}  // namespace plum
