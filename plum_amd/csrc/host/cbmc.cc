// Grand-canonical chain exchange with configurational bias: the growth,
// insertion, deletion and Widom test insertion of the reference
// (src/force_field/cbmc.cc:152-533) on std::mt19937 with the reference's
// order of draws.  A whole growth runs on the GPU in one call
// (plum_cbmc_grow); CBMCFGenTrialBeads, the one-step surface, evaluates its
// K trials in one call (plum_cbmc_trials).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "force_field.h"

namespace plum {

namespace {

double u01(std::mt19937& g) { return (double)g() / g.max(); }

// randSphere (misc.cc:95-109), Marsaglia 1972
void rand_sphere(double v[3], std::mt19937& g) {
  double s = 2, r1 = 0, r2 = 0;
  while (s > 1) {
    r1 = 1 - 2 * u01(g);
    r2 = 1 - 2 * u01(g);
    s = r1 * r1 + r2 * r2;
  }
  const double h = 2 * std::sqrt(1 - s);
  v[0] = r1 * h;
  v[1] = r2 * h;
  v[2] = 1 - 2 * s;
}

}  // namespace

// gasdev (misc.cc:111-133): Marsaglia's polar method; the second deviate of
// each pair is kept for the next call (process-wide statics in the
// reference, one cache per ForceField here).
double ForceField::gasdev(double mean, double stdev, std::mt19937& g) {
  if (gauss_iset_ == 0) {
    double v1, v2, rsq;
    do {
      v1 = 2.0 * u01(g) - 1.0;
      v2 = 2.0 * u01(g) - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
    gauss_gset_ = v1 * fac;
    gauss_iset_ = 1;
    return v2 * fac * stdev + mean;
  }
  gauss_iset_ = 0;
  return gauss_gset_ * stdev + mean;
}

// PotentialSpring::RandomBondLen (potential_spring.cc:50-61)
double ForceField::random_bond_len(std::mt19937& g) {
  const double sigma = std::sqrt(1 / (r_.beta * r_.k_bond));
  const double a = (r_.r0 + 3 * sigma) * (r_.r0 + 3 * sigma);
  double len = 0;
  bool ready = false;
  while (!ready) {
    len = gasdev(r_.r0, sigma, g);
    ready = u01(g) < (len * len / a);
  }
  return len;
}

// Energies of K trials (tri: K chain beads then K counterions) against the
// system less molecules [delete_id, delete_id + counterions] and the chain
// grown so far (chain_: beads then counterions).
void ForceField::trial_energies(int K, int cur, int delete_id, const double* tri, double* e) {
  const bool ch = q_gc_ != 0;
  std::vector<double> grown;
  if (cur > 0) {
    grown.assign(chain_.begin(), chain_.begin() + 3 * cur);
    if (ch) grown.insert(grown.end(), chain_.begin() + 3 * gc_len_, chain_.begin() + 3 * (gc_len_ + cur));
  }
  plum_cbmc_request rq{};
  rq.n_trials = K;
  rq.grown = cur;
  rq.exclude_lo = 0;
  rq.exclude_hi = -1;
  if (delete_id >= 0) {
    const int last = delete_id + (ch ? gc_len_ : 0);
    rq.exclude_lo = offset_[delete_id];
    rq.exclude_hi = offset_[last + 1] - 1;
  }
  rq.type_chain = gc_type_;
  rq.type_ion = gc_type_;
  rq.charged = ch ? 1 : 0;
  rq.q_chain = q_gc_;
  rq.q_ion = -q_gc_;
  check(plum_cbmc_trials(ctx_, &rq, tri, grown.empty() ? nullptr : grown.data(), e), "BeadsEnergy");
}

// cbmc.cc:161-212.  The trials are drawn first, in the reference's order
// (bond length, direction, counterion), then evaluated together.
double ForceField::CBMCFGenTrialBeads(const Bead& end_bead, std::vector<Molecule>& mols, int current_len,
                                      std::mt19937& g, int delete_id) {
  sync(mols);
  const int K = gc_trials_;
  const bool ch = q_gc_ != 0;
  tri_.assign(6 * (size_t)K, 0.0);
  wts_.assign(K, 0.0);
  en_.assign(K, 0.0);
  for (int i = 0; i < K; i++) {
    double bond_len = 0;
    if (r_.use_bond) bond_len = random_bond_len(g);
    else if (r_.use_rigid_bond) bond_len = r_.rigid_bond;
    double v[3];
    rand_sphere(v, g);
    for (int j = 0; j < 3; j++) tri_[3 * i + j] = v[j] * bond_len + end_bead.GetCrd(0, j);
    if (ch)
      for (int j = 0; j < 3; j++) tri_[3 * (K + i) + j] = u01(g) * box_[j];
  }
  trial_energies(K, current_len, delete_id, tri_.data(), en_.data());
  double Wi = 0;
  for (int i = 0; i < K; i++) {
    wts_[i] = std::exp(-r_.beta * en_[i]);
    Wi += wts_[i];
  }
  return Wi;
}

// The draws of one growth in the reference's order: the first bead and its
// counterion (insertion, test insertion), then for each step s = 1..last the
// K trials (bond length, direction, counterion; CBMCFGenTrialBeads,
// cbmc.cc:161-212) and, for an insertion, the uniform of the choice among
// them (cbmc.cc:241) unless s == last and !sel_last.
void ForceField::draw_growth(std::mt19937& g, int mode, int last, bool sel_last, bool draw_first) {
  const int K = gc_trials_, L = gc_len_;
  const bool ch = q_gc_ != 0;
  if (draw_first) first_.assign(6, 0.0);
  disp_.assign(3 * (size_t)L * K, 0.0);
  ion_.assign(3 * (size_t)L * K, 0.0);
  sel_.assign(L, 0.0);
  if (mode != 1 && draw_first)
    for (int i = 0; i < (ch ? 2 : 1); i++)
      for (int j = 0; j < 3; j++) first_[3 * i + j] = u01(g) * box_[j];
  for (int s = 1; s <= last; s++) {
    for (int t = 0; t < K; t++) {
      double bond_len = 0;
      if (r_.use_bond) bond_len = random_bond_len(g);
      else if (r_.use_rigid_bond) bond_len = r_.rigid_bond;
      double v[3];
      rand_sphere(v, g);
      const size_t o = 3 * ((size_t)s * K + t);
      for (int j = 0; j < 3; j++) disp_[o + j] = v[j] * bond_len;
      if (ch)
        for (int j = 0; j < 3; j++) ion_[o + j] = u01(g) * box_[j];
    }
    if (mode != 1 && (s < last || sel_last)) sel_[s] = u01(g);
  }
}

// One whole growth on the GPU (plum_cbmc_grow); returns the Rosenbluth
// weight.  The draws are made up front; when the weight reaches 0 at step s
// the reference stops drawing there, so the stream is rewound and redrawn up
// to that point.
double ForceField::grow_device(std::vector<Molecule>& mols, std::mt19937& g, int mode, int delete_id,
                               bool draw_first) {
  sync(mols);
  const int L = gc_len_, K = gc_trials_;
  const bool ch = q_gc_ != 0;
  const std::mt19937 g0 = g;
  const int iset0 = gauss_iset_;
  const double gset0 = gauss_gset_;
  draw_growth(g, mode, L - 1, true, draw_first);
  plum_growth_request rq{};
  rq.mode = mode;
  rq.chain_len = L;
  rq.n_trials = K;
  rq.exclude_lo = 0;
  rq.exclude_hi = -1;
  if (mode == 1) {
    act_.assign(6 * (size_t)L, 0.0);
    for (int i = 0; i < L; i++)
      for (int j = 0; j < 3; j++) {
        act_[3 * (size_t)i + j] = mols[delete_id].bds[i].GetCrd(0, j);
        if (ch) act_[3 * (size_t)(L + i) + j] = mols[delete_id + 1 + i].bds[0].GetCrd(0, j);
      }
    rq.exclude_lo = offset_[delete_id];
    rq.exclude_hi = offset_[delete_id + (ch ? L : 0) + 1] - 1;
  }
  rq.type_chain = gc_type_;
  rq.type_ion = gc_type_;
  rq.charged = ch ? 1 : 0;
  rq.q_chain = q_gc_;
  rq.q_ion = -q_gc_;
  chain_.assign(6 * (size_t)L, 0.0);
  double W = 0;
  int stop = -1;
  check(plum_cbmc_grow(ctx_, &rq, first_.data(), disp_.data(), ion_.data(), sel_.data(),
                       mode == 1 ? act_.data() : nullptr, &W, &stop, chain_.data()),
        "CBMC growth");
  if (stop >= 0) {
    g = g0;
    gauss_iset_ = iset0;
    gauss_gset_ = gset0;
    draw_growth(g, mode, stop, false, draw_first);
  }
  return W;
}

// spc1: exchanged molecules present; spc2: their counterion species
// (cbmc.cc:271-279)
void ForceField::gc_counts(std::vector<Molecule>& mols, int& spc1, int& spc2) {
  UpdateMolCounts(mols);
  if (gc_len_ > 1) spc1 = n_chain_;
  else if (q_gc_ >= 0) spc1 = n_cion_ - r_.coion;
  else spc1 = n_aion_ - r_.coion;
  if (q_gc_ == 0) spc2 = 0;
  else if (q_gc_ > 0) spc2 = n_aion_;
  else spc2 = n_cion_;
}

// cbmc.cc:268-300: the insertion's factor (V / Lambda)^n / (N + 1)! ...
double ForceField::insertion_factor(std::vector<Molecule>& mols) {
  const int L = gc_len_;
  const bool ch = q_gc_ != 0;
  int spc1, spc2;
  gc_counts(mols, spc1, spc2);
  const int spc1_add = 1, spc2_add = ch ? L : 0;
  double factorial = 1;
  for (int i = spc1 + 1; i <= spc1 + spc1_add; i++) factorial *= i;
  for (int i = spc2 + 1; i <= spc2 + spc2_add; i++) factorial *= i;
  const double vol = box_[0] * box_[1] * box_[2];
  double vol_over_lam = 1;
  vol_over_lam *= std::pow(vol / (r_.debroglie / std::pow((double)L, 1.5)), spc1_add);
  vol_over_lam *= std::pow(vol / (r_.debroglie / std::pow(1.0, 1.5)), spc2_add);
  return vol_over_lam * (1.0 / factorial);
}

// cbmc.cc:403-441: the deletion's factor (Lambda / V)^n N! ...
double ForceField::deletion_factor(std::vector<Molecule>& mols) {
  const int L = gc_len_;
  const bool ch = q_gc_ != 0;
  int spc1, spc2;
  gc_counts(mols, spc1, spc2);
  const int spc1_del = 1, spc2_del = ch ? L : 0;
  double factorial = 1;
  for (int i = spc1; i > spc1 - spc1_del; i--) factorial *= i;
  for (int i = spc2; i > spc2 - spc2_del; i--) factorial *= i;
  const double vol = box_[0] * box_[1] * box_[2];
  double lam_over_vol = 1;
  lam_over_vol *= std::pow((r_.debroglie / std::pow((double)L, 1.5)) / vol, spc1_del);
  lam_over_vol *= std::pow((r_.debroglie / std::pow(1.0, 1.5)) / vol, spc2_del);
  return lam_over_vol * factorial;
}

// cbmc.cc:309-330: the grown chain (linear bonds) then one molecule per
// counterion, appended to mols.
void ForceField::append_chain(std::vector<Molecule>& mols) {
  const int L = gc_len_;
  const bool ch = q_gc_ != 0;
  Molecule m;
  for (int i = 0; i < L; i++) {
    Bead b(r_.gc_bead_symbol, -1, -1, q_gc_, chain_[3 * i], chain_[3 * i + 1], chain_[3 * i + 2]);
    b.SetType(gc_type_);
    m.AddBead(b);
  }
  for (int i = 0; i < L - 1; i++) m.AddBond(i, i + 1);
  mols.push_back(m);
  if (ch)
    for (int i = L; i < 2 * L; i++) {
      Molecule c;
      Bead b(r_.gc_bead_symbol, -1, -1, -q_gc_, chain_[3 * i], chain_[3 * i + 1], chain_[3 * i + 2]);
      b.SetType(gc_type_);
      c.AddBead(b);
      mols.push_back(c);
    }
}

// Boltzmann factor of a first chain bead and its counterion (first[0..5])
// against the system less molecules [delete_id, delete_id + counterions]
// (BeadsEnergy with nothing grown yet, one plum_cbmc_trials trial).
double ForceField::first_bead_weight(const double first[6], int delete_id) {
  double e = 0;
  trial_energies(1, 0, delete_id, first, &e);
  return std::exp(-r_.beta * e);
}

// One growth's ln(W / w_1): its sum and the running ln of the sum of W / w_1.
void ForceField::GCPrescreenStats::record(double log_rest) {
  sum_log_rest += log_rest;
  if (n_rest == 0) lse_rest = log_rest;
  else {
    const double hi = std::max(lse_rest, log_rest), lo = std::min(lse_rest, log_rest);
    lse_rest = hi + std::log1p(std::exp(lo - hi));
  }
  n_rest++;
}

void ForceField::SetGCPrescreen(int mode, double log_theta) {
  if (mode < 0 || mode > 2) throw std::runtime_error("SetGCPrescreen: mode must be 0, 1 or 2");
  gc_pre_mode_ = mode;
  gc_log_theta_ = log_theta;
}

bool ForceField::CBMCFChainInsertion(std::vector<Molecule>& mols, std::mt19937& g) {
  if (!r_.use_gc) throw std::runtime_error("CBMCFChainInsertion without grand-canonical moves");
  if (gc_pre_mode_ == 2) return insertion_two_stage(mols, g);
  sync(mols);
  const double weight = grow_device(mols, g, 0, -1);
  if (weight <= 0) return false;
  if (gc_pre_mode_ == 1) {
    // the statistic of the two-stage first stage (no draw: the chain is the
    // reference's)
    const double w1 = first_bead_weight(first_.data(), -1);
    pre_.record(std::log(weight / w1));
  }
  const double C = insertion_factor(mols);
  const double rn = u01(g);
  if (!(rn < (std::exp(r_.beta * r_.mu) * weight) * C)) return false;
  append_chain(mols);
  return true;
}

// Two-stage insertion: the first bead and counterion drawn as the growth
// draws them, then the first-stage uniform, then (on passing) the rest of the
// growth's draws and the second-stage uniform.
bool ForceField::insertion_two_stage(std::vector<Molecule>& mols, std::mt19937& g) {
  sync(mols);
  const bool ch = q_gc_ != 0;
  first_.assign(6, 0.0);
  for (int i = 0; i < (ch ? 2 : 1); i++)
    for (int j = 0; j < 3; j++) first_[3 * i + j] = u01(g) * box_[j];
  const double w1 = first_bead_weight(first_.data(), -1);
  if (w1 <= 0) {   // the reference's growth ends here as well
    pre_.s1_rejected[0]++;
    return false;
  }
  const double C = insertion_factor(mols);
  const double r1 = std::exp(r_.beta * r_.mu) * C * w1 * std::exp(gc_log_theta_);
  if (!(u01(g) < r1)) {
    pre_.s1_rejected[0]++;
    return false;
  }
  pre_.s1_passed[0]++;
  const double weight = grow_device(mols, g, 0, -1, false);
  if (weight <= 0) return false;
  pre_.record(std::log(weight / w1));
  // min(1, r / r_1) with r = e^{beta mu} C W
  if (!(u01(g) < weight / (w1 * std::exp(gc_log_theta_)))) return false;
  append_chain(mols);
  return true;
}

int ForceField::CBMCFChainDeletion(std::vector<Molecule>& mols, std::mt19937& g) {
  if (!r_.use_gc) throw std::runtime_error("CBMCFChainDeletion without grand-canonical moves");
  sync(mols);
  const int L = gc_len_;
  const bool ch = q_gc_ != 0;
  int spc1, spc2;
  gc_counts(mols, spc1, spc2);
  int del = (int)std::floor(u01(g) * spc1);
  if (del >= spc1) del = spc1 - 1;
  if (del < 0) return -1;  // nothing to delete (the reference indexes out of range)
  del = r_.phantom + r_.coion + r_.grafted + r_.grafted_counterion + del * (ch ? 1 + L : 1);
  const int counterion = ch ? L : 0;
  if (del + counterion >= (int)mols.size() || mols[del].Size() != L)
    throw std::runtime_error("CBMCFChainDeletion: molecules are not in the exchange layout (chain, then its counterions)");
  if (gc_pre_mode_ == 2) return deletion_two_stage(mols, g, del);
  const double weight = grow_device(mols, g, 1, del);
  const double C = deletion_factor(mols);
  const double rn = u01(g);
  return rn < C / (std::exp(r_.beta * r_.mu) * weight) ? del : -1;
}

// Two-stage deletion of chain `del` (chosen by the reference's draw): the
// first-stage uniform on the inverse of the insertion's first stage, the
// regrowth along the chain on passing, then the second-stage uniform.
int ForceField::deletion_two_stage(std::vector<Molecule>& mols, std::mt19937& g, int del) {
  const bool ch = q_gc_ != 0;
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 3; j++) {
    f[j] = mols[del].bds[0].GetCrd(0, j);
    if (ch) f[3 + j] = mols[del + 1].bds[0].GetCrd(0, j);
  }
  const double w1 = first_bead_weight(f, del);
  const double C = deletion_factor(mols);
  const double r1 = C / (std::exp(r_.beta * r_.mu) * w1 * std::exp(gc_log_theta_));
  if (!(u01(g) < r1)) {
    pre_.s1_rejected[1]++;
    return -1;
  }
  pre_.s1_passed[1]++;
  const double weight = grow_device(mols, g, 1, del);
  // min(1, r / r_1) with r = C / (e^{beta mu} W)
  return u01(g) < w1 * std::exp(gc_log_theta_) / weight ? del : -1;
}

double ForceField::CalcChemicalPotentialF(std::vector<Molecule>& mols, std::mt19937& g) {
  sync(mols);
  if (gc_len_ < 1) throw std::runtime_error("CalcChemicalPotentialF: no molecule species to insert");
  const int L = gc_len_, n_ins = 50;  // mu_tot_ins (force_field.cc:236)
  const bool ch = q_gc_ != 0;
  int spc1, spc2;
  gc_counts(mols, spc1, spc2);
  const int spc1_add = 1, spc2_add = ch ? L : 0;
  double factorial = 1;
  for (int i = spc1 + 1; i <= spc1 + spc1_add; i++) factorial *= i;
  for (int i = spc2 + 1; i <= spc2 + spc2_add; i++) factorial *= i;
  const double vol = box_[0] * box_[1] * box_[2];
  double vol_over_lam = 1;
  vol_over_lam *= std::pow(vol / 15.625 / (r_.debroglie / std::pow((double)L, 1.5)), spc1_add);
  vol_over_lam *= std::pow(vol / 15.625 / (r_.debroglie / std::pow(1.0, 1.5)), spc2_add);
  const double C = vol_over_lam * (1.0 / factorial);
  double total = 0;
  for (int c = 0; c < n_ins; c++) total += grow_device(mols, g, 2, -1) * C;
  return total / (double)n_ins;
}

}  // namespace plum
