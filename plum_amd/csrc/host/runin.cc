#include "runin.h"

#include <fstream>
#include <sstream>

namespace plum {

namespace {

// `name value` with the name ignored, as every reference reader does.
template <typename T>
bool next(std::istream& in, T& v, const char* what, std::string& err) {
  std::string flag;
  if (!(in >> flag >> v)) {
    err = std::string("run.in: missing or malformed value for ") + what;
    return false;
  }
  return true;
}

bool next_bool(std::istream& in, int& v, const char* what, std::string& err) {
  std::string flag, tok;
  if (!(in >> flag >> tok)) {
    err = std::string("run.in: missing value for ") + what;
    return false;
  }
  if (tok == "0" || tok == "1") {
    v = tok == "1";
    return true;
  }
  err = std::string("run.in: expected 0/1 for ") + what + ", got " + tok;
  return false;
}

// {flag symbol flag sigma flag eps}... until the symbol "end"
// (potential_truncated_lj.cc:29-46, potential_truncated_lj_wall.cc:24-38)
bool read_type_table(std::istream& in, std::vector<TypeParam>& out, const char* what, std::string& err) {
  for (;;) {
    std::string flag, symbol;
    if (!(in >> flag >> symbol)) {
      err = std::string("run.in: unterminated ") + what + " table (missing 'end')";
      return false;
    }
    if (symbol == "end") return true;
    TypeParam t;
    t.symbol = symbol;
    if (!(in >> flag >> t.sigma >> flag >> t.eps)) {
      err = std::string("run.in: malformed ") + what + " entry for " + symbol;
      return false;
    }
    out.push_back(t);
  }
}

}  // namespace

bool read_sim_block(std::istream& in, RunIn& r, std::string& err) {
  return next(in, r.crd_file, "crd file", err) && next(in, r.top_file, "top file", err) &&
         next(in, r.out_prefix, "output prefix", err) && next(in, r.steps, "steps", err) &&
         next(in, r.steps_eq, "equilibrium steps", err) && next(in, r.sample_freq, "sample freq", err) &&
         next(in, r.stat_out_freq, "stat freq", err) && next(in, r.traj_out_freq, "traj freq", err) &&
         next(in, r.npbc, "PBC dimensions", err) && next(in, r.beta, "beta", err) &&
         next(in, r.move_size, "move size", err) && next(in, r.calc_chem_pot, "calc_chem_pot", err) &&
         next(in, r.phantom, "phantom", err) && next(in, r.coion, "coion", err) &&
         next(in, r.grafted, "grafted", err) && next(in, r.grafted_counterion, "grafted counterion", err) &&
         next(in, r.move_prob[0], "p(bead)", err) && next(in, r.move_prob[1], "p(COM)", err) &&
         next(in, r.move_prob[2], "p(pivot)", err) && next(in, r.move_prob[3], "p(crankshaft)", err) &&
         next(in, r.move_prob[4], "p(reptation)", err);
}

bool read_ff_block(std::istream& in, RunIn& r, std::string& err) {
  if (!next(in, r.vp_el_res, "vp resolution", err) || !next(in, r.vp_bead_size, "vp bead size", err) ||
      !next_bool(in, r.use_pair, "use_pair", err) || !next_bool(in, r.use_ewald, "use_ewald", err) ||
      !next_bool(in, r.use_bond, "use_bond", err) || !next_bool(in, r.use_rigid_bond, "use_rigid_bond", err) ||
      !next_bool(in, r.use_angle, "use_angle", err) || !next_bool(in, r.use_dihed, "use_dihed", err) ||
      !next_bool(in, r.use_ext, "use_ext", err) || !next_bool(in, r.use_gc, "use_gc", err))
    return false;
  if (r.use_gc) {
    if (!next(in, r.mu, "mu", err) || !next(in, r.debroglie, "deBroglie", err) ||
        !next(in, r.gc_freq, "gc_freq", err) || !next(in, r.gc_chain_len, "gc_chain_len", err) ||
        !next(in, r.gc_bead_charge, "gc_bead_charge", err) || !next(in, r.gc_bead_symbol, "gc_bead_symbol", err) ||
        !next(in, r.cbmc_trials, "cbmc_trials", err))
      return false;
  }
  if (r.use_pair) {
    if (!next(in, r.pair_name, "pair potential", err)) return false;
    if (r.pair_name != "TruncatedLJ") {
      err = "run.in: pair potential '" + r.pair_name + "' is not supported (TruncatedLJ only)";
      return false;
    }
    if (!next(in, r.lj_cutoff, "LJ cutoff", err) || !read_type_table(in, r.lj_types, "LJ", err)) return false;
  }
  if (r.use_ewald) {
    if (!next(in, r.ewald_name, "Ewald potential", err)) return false;
    if (r.ewald_name != "Coul") {
      err = "run.in: undefined Ewald potential '" + r.ewald_name + "'";
      return false;
    }
    if (!next(in, r.lB, "Bjerrum length", err) || !next(in, r.dielectric, "dielectric", err) ||
        !next(in, r.alpha, "Ewald alpha", err) || !next_bool(in, r.dipole_correction, "dipole correction", err))
      return false;
  }
  if (r.use_bond) {
    if (!next(in, r.bond_name, "bond potential", err)) return false;
    if (r.bond_name != "Spring") {
      err = "run.in: undefined bonded potential '" + r.bond_name + "'";
      return false;
    }
    if (!next(in, r.k_bond, "spring k", err) || !next(in, r.r0, "spring r0", err)) return false;
  }
  if (r.use_rigid_bond && !next(in, r.rigid_bond, "rigid bond length", err)) return false;
  if (r.use_ext) {
    if (!next(in, r.ext_name, "external potential", err)) return false;
    if (r.ext_name != "TruncatedLJWall") {
      err = "run.in: external potential '" + r.ext_name + "' is not supported (TruncatedLJWall only)";
      return false;
    }
    if (!next(in, r.wall_cut, "wall cutoff", err) || !next(in, r.wall_sigma, "wall sigma", err) ||
        !next(in, r.wall_eps, "wall epsilon", err) || !read_type_table(in, r.wall_types, "wall", err))
      return false;
  }
  return true;
}

bool read_top(const std::string& path, int& n_bead, int& n_mol, double box[3], std::vector<Molecule>& mols,
              std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open topology file " + path;
    return false;
  }
  std::string flag;
  if (!(f >> flag >> n_bead >> flag >> n_mol >> flag >> box[0] >> flag >> box[1] >> flag >> box[2])) {
    err = "malformed topology header in " + path;
    return false;
  }
  if (n_mol < 0 || n_bead < 0) {
    err = "topology declares a negative bead or molecule count";
    return false;
  }
  mols.assign(n_mol, Molecule());
  // bonds: `mol i j` until -1; angles and dihedrals are read and ignored
  const int arity[3] = {2, 3, 4};
  for (int sec = 0; sec < 3; sec++) {
    int mol = 0;
    if (!(f >> flag >> mol)) {
      err = "malformed topology section in " + path;
      return false;
    }
    while (mol != -1) {
      int ind[4] = {0, 0, 0, 0};
      for (int k = 0; k < arity[sec]; k++)
        if (!(f >> ind[k])) {
          err = "malformed topology entry in " + path;
          return false;
        }
      if (mol < 0 || mol >= n_mol) {
        err = "topology entry for molecule out of range";
        return false;
      }
      if (sec == 0) mols[mol].AddBond(ind[0], ind[1]);
      if (!(f >> mol)) {
        err = "topology section not terminated by -1";
        return false;
      }
    }
  }
  return true;
}

bool read_crd(const std::string& path, std::vector<Molecule>& mols, int& id_counter, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open coordinate file " + path;
    return false;
  }
  std::string line;
  std::getline(f, line);
  std::getline(f, line);
  int mol_id;
  std::string symbol;
  double x, y, z, q;
  while (f >> mol_id >> symbol >> x >> y >> z >> q) {
    // simulation.cc:594 accepts mol_id == n_mol (one past the end); that
    // index is rejected here
    if (mol_id < 0 || mol_id >= (int)mols.size()) {
      err = "molecule ID " + std::to_string(mol_id) + " is out of bounds";
      return false;
    }
    mols[mol_id].AddBead(Bead(symbol, id_counter++, mol_id, q, x, y, z));
  }
  if (!f.eof()) {
    err = "malformed coordinate line in " + path;
    return false;
  }
  for (size_t m = 0; m < mols.size(); m++)
    if (mols[m].Size() == 0) {
      err = "molecule " + std::to_string(m) + " has no beads in the coordinate file";
      return false;
    }
  return true;
}

}  // namespace plum
