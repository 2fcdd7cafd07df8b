#include "simulation.h"

#include <chrono>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace plum {

namespace {

// (double)rand_gen()/rand_gen.max(), the reference's uniform draw
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

std::string join(const std::string& dir, const std::string& f) {
  if (dir.empty() || (!f.empty() && f[0] == '/')) return f;
  return dir + "/" + f;
}

}  // namespace

// Molecule::BeadTranslate (molecule.cc:103-133): 3 move_size along a random
// direction from the current position of the first bead.
void BeadTranslate(Molecule& m, double move_size, std::mt19937& g) {
  double v[3];
  rand_sphere(v, g);
  double vl = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (vl > 0) vl = 3 * move_size / vl;
  m.bds[0].SetMoved();
  for (int i = 0; i < 3; i++) m.bds[0].SetCrd(1, i, m.bds[0].GetCrd(0, i) + vl * v[i]);
}

// Molecule::COMTranslate (molecule.cc:135-151)
void COMTranslate(Molecule& m, double move_size, std::mt19937& g) {
  double v[3];
  for (int k = 0; k < 3; k++) v[k] = 0.5 * move_size * u01(g);
  for (int k = 0; k < 3; k++)
    if (g() % 2 == 0) v[k] = -v[k];
  for (Bead& b : m.bds) {
    for (int k = 0; k < 3; k++) b.SetCrd(1, k, b.GetCrd(0, k) + v[k]);
    b.SetMoved();
  }
}

// Molecule::Pivot (molecule.cc:153-237): every bond beyond the pivot gets a
// random kick of size move_size*U, is renormalised to the (varied) bond
// length, and the chain beyond it follows.
void Pivot(Molecule& m, double move_size, std::mt19937& g, double eq, bool vary) {
  const int len = m.Size();
  int pivot = (int)std::floor(len * u01(g));
  if (pivot == len) pivot--;
  const double msr = move_size * u01(g);
  auto step = [&](int i, int nb, int dir) {
    double v[3];
    rand_sphere(v, g);
    double x[3], d[3];
    for (int k = 0; k < 3; k++) {
      x[k] = m.bds[i].GetCrd(1, k) + msr * v[k];
      d[k] = x[k] - m.bds[nb].GetCrd(1, k);
    }
    double bl = eq;
    if (vary) bl += (eq / 5.0) * (u01(g) - 0.5);
    const double norm = bl / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double mv[3];
    for (int k = 0; k < 3; k++) mv[k] = norm * d[k] + m.bds[nb].GetCrd(1, k) - m.bds[i].GetCrd(1, k);
    for (int j = i; j >= 0 && j < len; j += dir)
      for (int k = 0; k < 3; k++) m.bds[j].SetCrd(1, k, m.bds[j].GetCrd(1, k) + mv[k]);
  };
  for (int i = pivot + 1; i < len; i++) step(i, i - 1, 1);
  for (int i = pivot - 1; i >= 0; i--) step(i, i + 1, -1);
  for (Bead& b : m.bds) b.SetMoved();
}

// Molecule::Crankshaft (molecule.cc:239-265); the axis-angle rotation is the
// Rodrigues matrix c I + s [a]x + (1-c) a a^T.
void Crankshaft(Molecule& m, double move_size, std::mt19937& g) {
  const int len = m.Size();
  const int first = (int)std::floor((len - 2) * u01(g));
  int last = (int)std::floor((len - first) * u01(g)) + first;
  if (last >= len) last = len - 1;  // the reference indexes past the end when mt() == max
  double ax[3];
  for (int i = 0; i < 3; i++) ax[i] = m.bds[last].GetCrd(0, i) - m.bds[first].GetCrd(0, i);
  const double nn = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
  if (nn > 0) {
    const double inv = std::sqrt(nn);
    for (int i = 0; i < 3; i++) ax[i] /= inv;
  }
  const double angle = move_size * (u01(g) * 2 * M_PI - M_PI);
  const double sn = std::sin(angle), c = std::cos(angle);
  double R[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = (1 - c) * ax[i] * ax[j] + (i == j ? c : 0);
  R[0][1] -= sn * ax[2]; R[1][0] += sn * ax[2];
  R[0][2] += sn * ax[1]; R[2][0] -= sn * ax[1];
  R[1][2] -= sn * ax[0]; R[2][1] += sn * ax[0];
  for (int i = first + 1; i < last; i++) {
    m.bds[i].SetMoved();
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = m.bds[i].GetCrd(0, k) - m.bds[first].GetCrd(0, k);
    for (int k = 0; k < 3; k++)
      m.bds[i].SetCrd(1, k, R[k][0] * v[0] + R[k][1] * v[1] + R[k][2] * v[2] + m.bds[first].GetCrd(0, k));
  }
}

// Molecule::RandomReptation (molecule.cc:267-312)
void RandomReptation(Molecule& m, std::mt19937& g, double eq, bool vary) {
  if (eq < 0) throw std::runtime_error("reptation needs a bond length (rigid bond or spring)");
  double bl = eq;
  if (vary) bl += (eq / 5.0) * (u01(g) - 0.5);
  const int len = m.Size();
  int dir = 1, begin = 0, end = len - 1;
  if (g() % 2 == 0) {
    dir = -1;
    begin = len - 1;
    end = 0;
  }
  for (int i = begin; i != end; i += dir)
    for (int k = 0; k < 3; k++) m.bds[i].SetCrd(1, k, m.bds[i + dir].GetCrd(0, k));
  double v[3];
  rand_sphere(v, g);
  const double vl = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; k++) m.bds[end].SetCrd(1, k, m.bds[end].GetCrd(0, k) + bl * v[k] / vl);
  for (Bead& b : m.bds) b.SetMoved();
}

Simulation::Simulation(std::istream& in, unsigned seed, const std::string& base_dir) : rng(seed) {
  std::string err;
  if (!read_sim_block(in, r, err)) throw std::runtime_error(err);
  if (r.npbc < 2 || r.npbc > 3) throw std::runtime_error("invalid number of periodic dimensions");
  if (r.beta <= 0) throw std::runtime_error("beta must be > 0");
  double tot = 0;
  for (double p : r.move_prob) tot += p;
  if (std::fabs(tot - 1.0) > 1e-4) throw std::runtime_error("MC move probabilities do not sum up to 1.0");
  int n_bead = 0, n_mol = 0, id_counter = 0;
  if (!read_top(join(base_dir, r.top_file), n_bead, n_mol, box, mols, err)) throw std::runtime_error(err);
  if (!read_crd(join(base_dir, r.crd_file), mols, id_counter, err)) throw std::runtime_error(err);
  RunIn ffr = r;
  if (!read_ff_block(in, ffr, err)) throw std::runtime_error(err);
  r = ffr;
  if (r.move_prob[4] > 0 && !r.use_rigid_bond)
    throw std::runtime_error("reptation requires the rigid-bond setting (simulation.cc:182-186)");
  ff.Initialize(r, box, mols);
  update_counts();
}

// Simulation::UpdateMolCounts (simulation.cc:1066-1094)
void Simulation::update_counts() {
  n_mol = (int)mols.size();
  n_chain = n_cion = n_aion = n_nion = 0;
  for (int i = r.phantom; i < (int)mols.size(); i++) {
    if (mols[i].Size() > 1) n_chain++;
    else if (mols[i].bds[0].Charge() > 0) n_cion++;
    else if (mols[i].bds[0].Charge() < 0) n_aion++;
    else n_nion++;
  }
  n_chain -= r.grafted;
}

// Simulation::Run (simulation.cc:216-239) without output.
void Simulation::Run(long long steps) {
  update_counts();
  for (long long s = 0; s < steps; s++) {
    const int rn = (int)rng();  // the reference keeps the draw in an int
    if (ff.UseGC() && rn % r.gc_freq == 0) GCMove();
    else TranslationalMove();
    steps_done++;
  }
}

// n translational moves as one batch of the engine's device loop.
void Simulation::device_moves(long long n) {
  if (n == 0 || (int)mols.size() - r.phantom <= 0) return;
  plum_mc_params mc{};
  for (int k = 0; k < 5; k++) mc.move_prob[k] = r.move_prob[k];
  mc.move_size = r.move_size;
  mc.phantom = r.phantom;
  mc.grafted = r.grafted;
  mc.bond_len = 0;
  mc.vary_bond = 0;
  if (ff.UseBondRigid()) mc.bond_len = ff.RigidBondLen();
  if (ff.UseBondPot()) {
    mc.bond_len = ff.EqBondLen();
    mc.vary_bond = 1;
  }
  mc.seed = ((unsigned long long)rng() << 32) | rng();
  const auto t0 = std::chrono::steady_clock::now();
  const plum_mc_stats st = ff.DeviceMoves(mols, mc, n);
  seconds_device += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int k = 0; k < 5; k++) {
    attempted[k] += st.attempted[k];
    accepted[k] += st.accepted[k];
  }
}

void Simulation::RunDevice(long long steps) {
  update_counts();
  long long pending = 0;
  for (long long s = 0; s < steps; s++) {
    const int rn = (int)rng();
    if (ff.UseGC() && rn % r.gc_freq == 0) {
      device_moves(pending);
      pending = 0;
      const auto t0 = std::chrono::steady_clock::now();
      GCMove();
      seconds_gc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else {
      pending++;
    }
    steps_done++;
  }
  device_moves(pending);
}

void Simulation::RunWithOutput(long long steps, const std::string& prefix, bool device) {
  update_counts();
  std::ofstream stat(prefix + "_stat.dat");
  std::ofstream tp(prefix + "_traj_p.xyz"), tc(prefix + "_traj_c.xyz"), ta(prefix + "_traj_a.xyz"),
      tn(prefix + "_traj_n.xyz");
  if (!stat || !tp || !tc || !ta || !tn) throw std::runtime_error("cannot write the output files of " + prefix);
  std::ostream* traj[4] = {&tp, &tc, &ta, &tn};
  if (avg.rho_z.empty()) avg.rho_z.assign(3 * (size_t)std::ceil(box[2] / avg.rho_z_res), 0.0);
  print_stat_header(stat);
  long long pending = 0;
  for (long long s = 0; s < steps; s++) {
    const long long step = steps_done + 1;
    const int rn = (int)rng();
    const bool gc = ff.UseGC() && rn % r.gc_freq == 0;
    // what Sample or a writer reads at the end of this step must be the
    // configuration after it: batched moves are flushed before those steps
    const bool post_eq = step > r.steps_eq;
    const bool reads = (post_eq && step % r.sample_freq == 0) || step % r.stat_out_freq == 0 ||
                       (post_eq && step % r.traj_out_freq == 0);
    if (gc) {
      device_moves(pending);
      pending = 0;
      const long long ga[2] = {gc_attempted[0], gc_attempted[1]}, gk[2] = {gc_accepted[0], gc_accepted[1]};
      GCMove();
      if (post_eq)
        for (int k = 0; k < 2; k++) {
          avg.gc_attempted[k] += gc_attempted[k] - ga[k];
          avg.gc_accepted[k] += gc_accepted[k] - gk[k];
        }
    } else if (device) {
      pending++;
    } else {
      TranslationalMove();
    }
    steps_done++;
    if (reads) {
      device_moves(pending);
      pending = 0;
    }
    sample(step);
    if (step % r.stat_out_freq == 0) print_stat(stat, step);
    if (post_eq && step % r.traj_out_freq == 0) print_traj(traj, step);
    if (post_eq && step % r.stat_out_freq == 0) {
      PrintLastCrd(prefix);
      PrintLastCrd(prefix, "2");
      PrintLastTop(prefix);
      print_last_rho_z(prefix);
    }
  }
  device_moves(pending);
}

// Simulation::Sample (simulation.cc:723-798); with walls the wall-force slit
// pressure (CalcPressureForceLJELSlit, pressure.cc:404-487) every
// sample_freq * 10 steps, without them the volume-perturbation pressure.
void Simulation::sample(long long step) {
  if (step > r.steps_eq && step % r.sample_freq == 0) {
    avg.ff_count++;
    if (ff.UsePairPot()) avg.pair += ff.TotPairEnergy();
    if (ff.UseEwaldPot()) avg.ewald += ff.TotEwaldEnergy();
    if (ff.UseBondPot()) avg.bond += ff.TotBondEnergy();
    if (ff.UseExtPot()) avg.ext += ff.TotExtEnergy();
    avg.density += (double)mols.size() / (box[0] * box[1] * box[2]);
    // Simulation::CalcRhoZ (simulation.cc:1029-1045)
    const int nb = (int)(avg.rho_z.size() / 3);
    for (int i = r.phantom; i < (int)mols.size(); i++) {
      const int id = mols[i].Size() > 1 ? 0 : (mols[i].bds[0].Charge() >= 0 ? 1 : 2);
      for (const Bead& b : mols[i].bds) {
        double z = b.GetCrd(0, 2);
        z -= box[2] * std::floor(z / box[2]);
        const int bin = (int)std::floor(z / avg.rho_z_res);
        if (bin >= 0 && bin < nb) avg.rho_z[3 * (size_t)bin + id] += 1.0 / (box[0] * box[1] * avg.rho_z_res);
      }
    }
    if (ff.UseExtPot() && step % (r.sample_freq * 10) == 0) ff.CalcPressureForceLJELSlit(mols);
    else if (step % (r.sample_freq * 10) == 0) ff.CalcPressureVolScalingHSELSlit(mols);
    if (r.calc_chem_pot) avg.chem_pot += ff.CalcChemicalPotentialF(mols, rng);
  }
  if (step > r.steps_eq && step % r.sample_freq == 0 && !mols.empty()) {
    avg.mol_count++;
    double xyz[3];
    RadiusOfGyrationXYZ(xyz);
    avg.rg += RadiusOfGyration();
    for (int k = 0; k < 3; k++) avg.rg_xyz[k] += xyz[k];
    avg.ete += EndToEndDistance();
  }
}

// Simulation::PrintStatHeader (simulation.cc:685-721)
void Simulation::print_stat_header(std::ostream& o) const {
  o << "#Step";
  if (ff.UsePairPot()) o << " <PairEne>";
  if (ff.UseEwaldPot()) o << " <EwaldEne>";
  if (ff.UseBondPot()) o << " <BondEne>";
  if (ff.UseExtPot()) o << " <ExtPotEne>";
  o << " <TotalEnergy>";
  if (ff.UseGC()) o << " NoOfMol <Density>";
  if (ff.UseExtPot()) o << " <Pzz_LJ_ion> <Pzz_LJ_pol> <Pzz_LJ_wal> <Pzz_EL_ion> <Pzz_EL_pol> <Pzz_EL_wal>";
  else o << " <P>";
  if (r.calc_chem_pot) o << " mu";
  o << " <|Rg|> <|Rgx|> <|Rgy|> <|Rgz|> <ete>";
  o << " beadtrans comtrans pivot crankshaft reptate";
  if (ff.UseGC()) o << " insert delete";
  o << '\n';
}

// Simulation::PrintStat (simulation.cc:800-867): running averages with the
// reference's stream precisions (7 digits for the energies, then 4).
void Simulation::print_stat(std::ostream& o, long long step) const {
  o << step;
  const double n = (double)avg.ff_count;
  double tot = 0;
  o.precision(7);
  if (ff.UsePairPot()) {
    o << ' ' << avg.pair / n;
    tot += avg.pair / n;
  }
  if (ff.UseEwaldPot()) {
    o << ' ' << avg.ewald / n;
    tot += avg.ewald / n;
  }
  if (ff.UseBondPot()) {
    o << ' ' << avg.bond / n;
    tot += avg.bond / n;
  }
  if (ff.UseExtPot()) {
    o << ' ' << avg.ext / n;
    tot += avg.ext / n;
  }
  o << ' ' << tot;
  if (ff.UseGC()) {
    o << ' ' << (int)mols.size();
    o.precision(4);
    o << ' ' << avg.density / n;
  }
  // ForceField::GetPressure (pressure.cc:490-500): the six slit-pressure
  // averages with walls, nan without them
  o << ' ' << ff.GetPressure();
  if (r.calc_chem_pot) {
    if (avg.chem_pot / n > 0) o << ' ' << -(1.0 / r.beta) * std::log(avg.chem_pot / n);
    else o << " INF";
  }
  const double m = (double)avg.mol_count;
  o.precision(4);
  o << ' ' << std::sqrt(avg.rg / m) << ' ' << std::sqrt(avg.rg_xyz[0] / m) << ' ' << std::sqrt(avg.rg_xyz[1] / m)
    << ' ' << std::sqrt(avg.rg_xyz[2] / m) << ' ' << std::sqrt(avg.ete / m);
  for (int k = 0; k < 5; k++) o << ' ' << accepted[k] / (double)attempted[k];
  if (ff.UseGC())
    o << ' ' << avg.gc_accepted[0] / (double)avg.gc_attempted[0] << ' '
      << avg.gc_accepted[1] / (double)avg.gc_attempted[1];
  o << '\n';
}

// Simulation::PrintTraj (simulation.cc:869-933): xyz frames of the chains,
// cations, anions and neutral ions, each molecule shifted by whole boxes so
// that its centre lies in the box.
void Simulation::print_traj(std::ostream* out[4], long long step) const {
  int n_chain_beads = 0;
  for (int i = r.phantom; i < (int)mols.size(); i++)
    if (mols[i].Size() > 1) n_chain_beads += mols[i].Size();
  const int counts[4] = {n_chain_beads, n_cion, n_aion, n_nion};
  const bool have[4] = {r.grafted + n_chain > 0, n_cion > 0, n_aion > 0, n_nion > 0};
  for (int f = 0; f < 4; f++)
    if (have[f]) *out[f] << counts[f] << "\nSTEP: " << step << '\n';
  for (int i = r.phantom; i < (int)mols.size(); i++) {
    const Molecule& m = mols[i];
    double c[3] = {0, 0, 0};
    for (const Bead& b : m.bds)
      for (int k = 0; k < 3; k++) c[k] += b.GetCrd(0, k);
    double disp[3];
    for (int k = 0; k < 3; k++) disp[k] = std::floor(c[k] / m.Size() / box[k]) * box[k];
    const int f = m.Size() > 1 ? 0 : (m.bds[0].Charge() > 0 ? 1 : (m.bds[0].Charge() < 0 ? 2 : 3));
    for (const Bead& b : m.bds)
      *out[f] << b.Symbol() << ' ' << b.GetCrd(0, 0) - disp[0] << ' ' << b.GetCrd(0, 1) - disp[1] << ' '
              << b.GetCrd(0, 2) - disp[2] << '\n';
  }
}

// Simulation::PrintLastRhoZ (simulation.cc:1047-1064)
void Simulation::print_last_rho_z(const std::string& prefix) const {
  const char* bck[2] = {"", "2"};
  const int nb = (int)(avg.rho_z.size() / 3);
  for (int f = 0; f < 2; f++) {
    std::ofstream o(prefix + "_lastrhoZ.dat" + bck[f]);
    if (!o) throw std::runtime_error("cannot write " + prefix + "_lastrhoZ.dat" + bck[f]);
    for (int bin = 0; bin < nb; bin++)
      o << (bin + 0.5) * avg.rho_z_res << '\t' << avg.rho_z[3 * (size_t)bin] / (double)avg.ff_count << '\t'
        << avg.rho_z[3 * (size_t)bin + 1] / (double)avg.ff_count << '\t'
        << avg.rho_z[3 * (size_t)bin + 2] / (double)avg.ff_count << '\n';
  }
}

// Simulation::GCMove (simulation.cc:357-404)
void Simulation::GCMove() {
  const bool insert = rng() % 2;
  if (insert) {
    gc_attempted[0]++;
    if (ff.CBMCFChainInsertion(mols, rng)) {
      gc_accepted[0]++;
      ff.EnergyInitForAddedMolecule(mols);
    }
  } else {
    gc_attempted[1]++;
    int deleted = -1;
    if (n_mol - r.phantom - r.coion > 0) deleted = ff.CBMCFChainDeletion(mols, rng);
    if (deleted >= r.phantom + r.coion) {
      gc_accepted[1]++;
      int ions = 0;
      for (const Bead& b : mols[deleted].bds) ions += (int)std::round(std::fabs(b.Charge()));
      mols.erase(mols.begin() + deleted, mols.begin() + deleted + ions + 1);
    }
  }
  update_counts();
  ff.UpdateMolCounts(mols);
}

// Simulation::TranslationalMove (simulation.cc:241-355)
void Simulation::TranslationalMove() {
  const int n_mol = (int)mols.size();
  if (n_mol - r.phantom <= 0) return;
  const int n_ions = n_cion + n_aion + n_nion;
  int which_chain = (int)std::floor(u01(rng) * (r.grafted + n_chain));
  int which_ion = (int)std::floor(u01(rng) * n_ions);
  if (which_chain == r.grafted + n_chain) which_chain--;
  if (which_ion == n_ions) which_ion--;
  int chain_id = -1, ion_id = -1, cc = -1, ic = -1;
  for (int i = r.phantom; i < n_mol; i++) {
    if (mols[i].Size() > 1) cc++;
    else ic++;
    if (cc == which_chain) { chain_id = i; which_chain = -2; }
    if (ic == which_ion) { ion_id = i; which_ion = -2; }
  }
  int mt = 0;
  const double rn = u01(rng);
  double cur = r.move_prob[0];
  while (cur < rn) {
    mt++;
    cur += r.move_prob[mt];
  }
  if (mt >= 5) mt = 4;
  int mol_id = -1;
  if (mt == 0 && n_ions > 0) {
    mol_id = ion_id;
    BeadTranslate(mols[mol_id], r.move_size, rng);
  } else if (r.grafted + n_chain > 0) {
    mol_id = chain_id;
    double bond_len = 0;
    bool vary = false;
    if (ff.UseBondRigid()) bond_len = ff.RigidBondLen();
    if (ff.UseBondPot()) {
      bond_len = ff.EqBondLen();
      vary = true;
    }
    switch (mt) {
      case 1: COMTranslate(mols[mol_id], r.move_size, rng); break;
      case 2: Pivot(mols[mol_id], r.move_size, rng, bond_len, vary); break;
      case 3: Crankshaft(mols[mol_id], r.move_size, rng); break;
      case 4: RandomReptation(mols[mol_id], rng, bond_len, vary); break;
    }
  }
  if ((mt == 0 && n_ions > 0) || (mt > 0 && r.grafted + n_chain > 0)) {
    attempted[mt]++;
    const double dE = ff.EnergyDifference(mols, mol_id);
    bool accept = false;
    if (dE < 1e8) accept = u01(rng) < std::exp(-r.beta * dE);
    ff.FinalizeEnergies(mols, accept, mol_id);
    for (Bead& b : mols[mol_id].bds) {
      if (accept) b.UpdateCurrentPos();
      else b.UpdateTrialPos();
      b.UnsetMoved();
    }
    if (accept) accepted[mt]++;
  }
}

int Simulation::DensityProfile(double res, std::vector<double>& out) const {
  if (!(res > 0)) throw std::runtime_error("density profile bin width must be > 0");
  const int nb = (int)std::floor(box[2] / res);
  out.assign(3 * (size_t)std::max(nb, 0), 0.0);
  for (int i = r.phantom; i < (int)mols.size(); i++) {
    const int sp = mols[i].Size() > 1 ? 0 : (mols[i].bds[0].Charge() >= 0 ? 1 : 2);
    for (const Bead& b : mols[i].bds) {
      double z = b.GetCrd(0, 2);
      z -= box[2] * std::floor(z / box[2]);
      const int bin = (int)std::floor(z / res);
      if (bin >= 0 && bin < nb) out[3 * (size_t)bin + sp] += 1.0 / (box[0] * box[1] * res);
    }
  }
  return nb;
}

void Simulation::PrintLastCrd(const std::string& prefix, const std::string& suffix) const {
  std::ofstream f(prefix + "_lastcrd.dat" + suffix);
  if (!f) throw std::runtime_error("cannot write " + prefix + "_lastcrd.dat" + suffix);
  int n = 0;
  for (const Molecule& m : mols) n += m.Size();
  f << n << "\n \n";
  for (int i = 0; i < (int)mols.size(); i++) {
    const Molecule& m = mols[i];
    double c[3] = {0, 0, 0}, shift[3] = {0, 0, 0};
    for (const Bead& b : m.bds)
      for (int k = 0; k < 3; k++) c[k] += b.GetCrd(0, k);
    if (i >= r.phantom)
      for (int k = 0; k < 3; k++) shift[k] = std::floor(c[k] / m.Size() / box[k]) * box[k];
    for (const Bead& b : m.bds) {
      f << i << ' ' << b.Symbol();
      f.precision(18);
      for (int k = 0; k < 3; k++) f << ' ' << b.GetCrd(0, k) - shift[k];
      f.precision(4);
      f << ' ' << b.Charge() << '\n';
    }
  }
}

void Simulation::PrintLastTop(const std::string& prefix) const {
  std::ofstream f(prefix + "_lasttop.dat");
  if (!f) throw std::runtime_error("cannot write " + prefix + "_lasttop.dat");
  int n = 0;
  for (const Molecule& m : mols) n += m.Size();
  f << "TotNoOfBeads: " << n << "\nTotNoOfMolec: " << mols.size() << '\n';
  f.precision(18);
  f << "Box_Length_X: " << box[0] << "\nBox_Length_Y: " << box[1] << "\nBox_Length_Z: " << box[2] << '\n';
  // the reference writes the bond, angle and dihedral lists empty
  f << "bonds:\n-1\nangles:\n-1\ndiheds\n-1\n";
}

// Simulation::RadiusOfGyration (simulation.cc:406-434)
double Simulation::RadiusOfGyration() const {
  double rg = 0;
  for (const Molecule& m : mols) {
    if (m.Size() <= 1) continue;
    double rm = 0;
    for (int j = 0; j < m.Size() - 1; j++)
      for (int k = j + 1; k < m.Size(); k++) {
        double s = 0;
        for (int e = 0; e < 3; e++) {
          const double d = m.bds[j].GetCrd(1, e) - m.bds[k].GetCrd(1, e);
          s += d * d;
        }
        rm += s;
      }
    rg += rm / (m.Size() * m.Size());
  }
  return (r.grafted + n_chain) > 0 ? rg / (r.grafted + n_chain) : 0;
}

// Simulation::RadiusOfGyrationXYZ (simulation.cc:437-481): per-axis
// absolute separations of the trial coordinates, no periodic folding.
void Simulation::RadiusOfGyrationXYZ(double out[3]) const {
  out[0] = out[1] = out[2] = 0;
  for (const Molecule& m : mols) {
    if (m.Size() <= 1) continue;
    double a[3] = {0, 0, 0};
    for (int j = 0; j < m.Size() - 1; j++)
      for (int k = j + 1; k < m.Size(); k++)
        for (int e = 0; e < 3; e++) {
          const double d = std::fabs(m.bds[j].GetCrd(1, e) - m.bds[k].GetCrd(1, e));
          a[e] += d * d;
        }
    for (int e = 0; e < 3; e++) out[e] += a[e] / (m.Size() * m.Size());
  }
  if (r.grafted + n_chain > 0)
    for (int e = 0; e < 3; e++) out[e] /= (r.grafted + n_chain);
}

// Simulation::EndToEndDistance (simulation.cc:483-501), mean squared
double Simulation::EndToEndDistance() const {
  double e = 0;
  for (const Molecule& m : mols) {
    if (m.Size() <= 1) continue;
    double s = 0;
    for (int k = 0; k < 3; k++) {
      const double d = m.bds[0].GetCrd(1, k) - m.bds[m.Size() - 1].GetCrd(1, k);
      s += d * d;
    }
    e += s;
  }
  return (r.grafted + n_chain) > 0 ? e / (r.grafted + n_chain) : 0;
}

}  // namespace plum
