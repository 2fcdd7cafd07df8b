#include "force_field.h"
#include "plum_host.h"

#include <cmath>
#include <sstream>
#include <stdexcept>

namespace plum {

ForceField::~ForceField() {
  if (ctx_) plum_destroy(ctx_);
}

void ForceField::check(int rc, const char* what) {
  if (rc != PLUM_OK)
    throw std::runtime_error(std::string("ForceField::") + what + ": " + plum_last_error());
}

int ForceField::TypeOf(const std::string& symbol) const {
  auto it = types_.find(symbol);
  return it == types_.end() ? -1 : it->second;
}

double ForceField::EqBondLen() const {
  if (!r_.use_bond) throw std::runtime_error("Illegal request of bond length (no bond potential)");
  return r_.r0;
}

void ForceField::Initialize(double beta, int npbc, double box_l[3], std::vector<Molecule>& mols, int phantom,
                            int coion, int grafted, int grafted_counterion, std::istream& in) {
  RunIn r;
  r.beta = beta;
  r.npbc = npbc;
  r.phantom = phantom;
  r.coion = coion;
  r.grafted = grafted;
  r.grafted_counterion = grafted_counterion;
  std::string err;
  if (!read_ff_block(in, r, err)) throw std::runtime_error(err);
  Initialize(r, box_l, mols);
}

void ForceField::build_params(plum_params& p, const double box_l[3]) {
  p = plum_params{};
  for (int k = 0; k < 3; k++) p.box[k] = box_l[k];
  p.npbc = r_.npbc;
  p.beta = r_.beta;
  p.use_pair = r_.use_pair;
  p.use_ewald = r_.use_ewald;
  p.use_bond = r_.use_bond;
  p.use_ext = r_.use_ext;
  p.lj_cutoff = r_.lj_cutoff;
  // the inert bead of an empty system has a type of its own with every
  // parameter zero
  p.n_types = (int)type_names_.size() + (ghost_ ? 1 : 0);
  // per-type tables; symbols missing from a table read 0, like the
  // reference's map<string,double>::operator[]
  for (size_t t = 0; t < type_names_.size(); t++) {
    const std::string& s = type_names_[t];
    for (const TypeParam& tp : r_.lj_types)
      if (tp.symbol == s) { p.lj_sigma[t] = tp.sigma; p.lj_eps[t] = tp.eps; }
    for (const TypeParam& tp : r_.wall_types)
      if (tp.symbol == s) { p.wall_sigma[t] = tp.sigma; p.wall_eps[t] = tp.eps; }
    p.type_kind[t] = s == "L" ? 1 : (s == "R" ? 2 : 0);
  }
  p.lB = r_.lB;
  p.dielectric = r_.dielectric;
  p.alpha = r_.alpha;
  p.dipole_correction = r_.dipole_correction;
  p.k_bond = r_.k_bond;
  p.r0 = r_.r0;
  p.wall_cut = r_.wall_cut;
  p.wall_sigma_w = r_.wall_sigma;
  p.wall_eps_w = r_.wall_eps;
  p.precision = PLUM_FP64;
}

void ForceField::assign_types(std::vector<Molecule>& mols) {
  auto add = [&](const std::string& s) {
    if (types_.count(s)) return;
    if ((int)type_names_.size() >= PLUM_MAX_TYPES - 1) throw std::runtime_error("too many bead types");
    types_[s] = (int)type_names_.size();
    type_names_.push_back(s);
  };
  for (const TypeParam& t : r_.lj_types) add(t.symbol);
  for (const TypeParam& t : r_.wall_types) add(t.symbol);
  if (r_.use_gc) add(r_.gc_bead_symbol);
  for (Molecule& m : mols)
    for (Bead& b : m.bds) {
      add(b.Symbol());
      b.SetType(types_[b.Symbol()]);
    }
}

void ForceField::Initialize(const RunIn& r, double box_l[3], std::vector<Molecule>& mols) {
  r_ = r;
  types_.clear();
  type_names_.clear();
  if (r_.use_gc) {
    if (r_.gc_freq <= 0) throw std::runtime_error("GCMC move frequency must be > 0");
    if (r_.gc_chain_len < 1 || r_.cbmc_trials < 1) throw std::runtime_error("CBMC chain length and trials must be >= 1");
    gc_len_ = r_.gc_chain_len;
    q_gc_ = r_.gc_bead_charge;
    gc_trials_ = r_.cbmc_trials;
  } else {
    // force_field.cc:243-255: the chemical potential of the first chain species
    gc_len_ = -1;
    for (int i = r_.phantom; i < (int)mols.size(); i++) {
      gc_len_ = mols[i].Size();
      q_gc_ = mols[i].bds[0].Charge();
      if (gc_len_ > 1) break;
    }
    r_.debroglie = 1;
    r_.gc_bead_symbol = "P";
    gc_trials_ = 30;
  }
  assign_types(mols);
  gc_type_ = TypeOf(r_.use_gc ? r_.gc_bead_symbol : "P");
  for (int k = 0; k < 3; k++) box_[k] = box_l[k];
  InitializeEnergy(mols);
  UpdateMolCounts(mols);
}

void ForceField::InitializeEnergy(std::vector<Molecule>& mols) {
  plum_params p;
  double box[3];
  offset_.assign(mols.size() + 1, 0);
  for (size_t m = 0; m < mols.size(); m++) offset_[m + 1] = offset_[m] + mols[m].Size();
  n_beads_ = offset_.back();
  ghost_ = n_beads_ == 0;
  const int n = ghost_ ? 1 : n_beads_;
  std::vector<double> xyz(3 * (size_t)n), q(n);
  std::vector<int> type(n), ms(ghost_ ? 2 : mols.size() + 1);
  for (size_t m = 0; m < mols.size(); m++) {
    ms[m] = offset_[m];
    for (int k = 0; k < mols[m].Size(); k++) {
      const Bead& b = mols[m].bds[k];
      const int i = offset_[m] + k;
      for (int e = 0; e < 3; e++) xyz[3 * (size_t)i + e] = b.GetCrd(0, e);
      q[i] = b.Charge();
      type[i] = b.Type();
    }
  }
  if (ghost_) {
    // no molecules: one uncharged bead with zero LJ and wall parameters at
    // the box centre stands in, so that the engine state exists
    for (int e = 0; e < 3; e++) xyz[e] = 0.5 * box_[e];
    q[0] = 0;
    type[0] = (int)type_names_.size();
    ms[0] = 0;
    ms[1] = 1;
  } else {
    ms[mols.size()] = n;
  }
  box[0] = box_[0];
  box[1] = box_[1];
  box[2] = box_[2];
  build_params(p, box);
  if (ctx_) {
    plum_destroy(ctx_);
    ctx_ = nullptr;
  }
  check(plum_create(&p, n, xyz.data(), q.data(), type.data(), (int)ms.size() - 1, ms.data(), &ctx_),
        "InitializeEnergy");
}

// Rebuilds the energy state when the molecule list no longer matches it
// (after a grand-canonical insertion or deletion).
void ForceField::sync(std::vector<Molecule>& mols) {
  int n = 0;
  for (const Molecule& m : mols) n += m.Size();
  if (offset_.size() == mols.size() + 1 && n == n_beads_) return;
  assign_types(mols);
  InitializeEnergy(mols);
}

void ForceField::EnergyInitForAddedMolecule(std::vector<Molecule>& mols) { sync(mols); }

double ForceField::EnergyDifference(std::vector<Molecule>& mols, int moved_mol) {
  sync(mols);
  const Molecule& mol = mols.at(moved_mol);
  std::vector<int> idx;
  std::vector<double> tr;
  idx.reserve(mol.Size());
  tr.reserve(3 * (size_t)mol.Size());
  for (int k = 0; k < mol.Size(); k++) {
    const Bead& b = mol.bds[k];
    if (!b.GetMoved()) continue;
    idx.push_back(offset_[moved_mol] + k);
    for (int e = 0; e < 3; e++) tr.push_back(b.GetCrd(1, e));
  }
  double total = 0;
  check(plum_energy_difference(ctx_, moved_mol, (int)idx.size(), idx.data(), tr.data(), &last_dE_, &total),
        "EnergyDifference");
  return total;
}

plum_mc_stats ForceField::DeviceMoves(std::vector<Molecule>& mols, const plum_mc_params& mc, long long n) {
  sync(mols);
  plum_mc_stats st{};
  check(plum_mc_run(ctx_, &mc, n, &st), "DeviceMoves");
  std::vector<double> xyz(3 * (size_t)std::max(1, n_beads_));
  check(plum_get_positions(ctx_, xyz.data()), "DeviceMoves");
  for (size_t m = 0; m < mols.size(); m++)
    for (int k = 0; k < mols[m].Size(); k++) {
      const double* p = &xyz[3 * (size_t)(offset_[m] + k)];
      mols[m].bds[k].SetAllCrd(p);
    }
  return st;
}

void ForceField::FinalizeEnergies(std::vector<Molecule>&, bool accept, int) {
  check(plum_finalize(ctx_, accept ? 1 : 0), "FinalizeEnergies");
}

plum_terms ForceField::Totals() {
  plum_terms t;
  check(plum_totals(ctx_, &t), "Totals");
  return t;
}

plum_terms ForceField::FullEnergy() {
  plum_terms t;
  check(plum_energy_full(ctx_, &t), "FullEnergy");
  return t;
}

plum_vp ForceField::CalcPressureVolScalingHSELSlit(std::vector<Molecule>& mols) {
  sync(mols);
  plum_vp v{};
  check(plum_vp_sample(ctx_, r_.phantom, &v), "CalcPressureVolScalingHSELSlit");
  vp_z_++;
  vp_du_ += v.dU;
  vp_ideal_ += v.n_mol / v.vol;
  const double dz = v.area_dz / (box_[0] * box_[1]);
  vp_dm_ += std::pow(1.0 + dz / box_[2], v.n_mol) * std::exp(-r_.beta * v.dU);
  vp_adz_ = v.area_dz;
  return v;
}

double ForceField::PressureYethiraj() const {
  if (vp_z_ == 0) return 0;
  return (r_.beta * vp_ideal_ - vp_du_ / vp_adz_) / vp_z_;
}

double ForceField::PressureDeMiguel() const {
  if (vp_z_ == 0) return 0;
  return r_.beta / vp_adz_ * std::log(vp_dm_ / vp_z_);
}

void ForceField::CalcPressureForceLJELSlit(std::vector<Molecule>& mols) {
  slit_z_++;
  std::vector<double> xyz;
  std::vector<double> q;
  std::vector<int> type, ms(1, 0);
  for (const Molecule& m : mols) {
    for (int k = 0; k < m.Size(); k++) {
      const Bead& b = m.bds[k];
      for (int e = 0; e < 3; e++) xyz.push_back(b.GetCrd(0, e));
      q.push_back(b.Charge());
      type.push_back(b.Type());
    }
    ms.push_back((int)q.size());
  }
  plum_params p{};
  double box[3] = {box_[0], box_[1], box_[2]};
  build_params(p, box);
  plum_info e{};
  check(plum_get_info(ctx_, &e), "CalcPressureForceLJELSlit");
  double f[6];
  check(plum_slit_forces(&p, &e, (int)q.size(), xyz.data(), q.data(), type.data(), (int)mols.size(), ms.data(),
                         r_.phantom, f),
        "CalcPressureForceLJELSlit");
  // pressure.cc:438-487: the wall-wall parts once (first call), the others
  // cumulative and averaged over the calls
  p_tensor_[6] += f[0];
  p_tensor_[7] += f[1];
  p_tensor_[9] += f[3];
  p_tensor_[10] += f[4];
  if (slit_z_ == 1) {
    p_tensor_[8] += f[2];
    p_tensor_[11] += f[5];
  }
  const double area = box_[0] * box_[1];
  if (slit_z_ == 1) {
    p_tensor_[2] = p_tensor_[8] / area;
    p_tensor_[5] = p_tensor_[11] / area;
  }
  p_tensor_[0] = p_tensor_[6] / (slit_z_ * area);
  p_tensor_[1] = p_tensor_[7] / (slit_z_ * area);
  p_tensor_[3] = p_tensor_[9] / (slit_z_ * area);
  p_tensor_[4] = p_tensor_[10] / (slit_z_ * area);
}

std::string ForceField::GetPressure() const {
  std::ostringstream foo;
  if (r_.use_ext)
    foo << p_tensor_[0] << " " << p_tensor_[1] << " " << p_tensor_[2] << " " << p_tensor_[3] << " " << p_tensor_[4]
        << " " << p_tensor_[5];
  else
    foo << "nan";
  return foo.str();
}

double ForceField::TotPairEnergy() { return Totals().lj; }
double ForceField::TotEwaldEnergy() {
  const plum_terms t = Totals();
  return t.real + t.kspace + t.self + t.dipole;
}
double ForceField::TotBondEnergy() { return Totals().bond; }
double ForceField::TotExtEnergy() { return Totals().wall; }

// force_field.cc:1258-1277
void ForceField::UpdateMolCounts(std::vector<Molecule>& mols) {
  sync(mols);
  n_chain_ = n_cion_ = n_aion_ = 0;
  for (int i = r_.phantom; i < (int)mols.size(); i++) {
    if (mols[i].Size() > 1) n_chain_++;
    else if (mols[i].bds[0].Charge() >= 0) n_cion_++;
    else n_aion_++;
  }
  n_chain_ -= r_.grafted;
}

}  // namespace plum
