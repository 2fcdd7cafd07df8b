// extern "C" surface of the host layer (include/plum_host.h).
#include <algorithm>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>

#include "plum_host.h"
#include "simulation.h"

struct plum_sim {
  plum::Simulation* s = nullptr;
};

namespace {

thread_local std::string g_host_err;

int fail(int code, const std::string& m) {
  g_host_err = m;
  return code;
}

std::string dir_of(const std::string& path) {
  const size_t p = path.find_last_of('/');
  return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

}  // namespace

extern "C" {

const char* plum_host_last_error(void) { return g_host_err.c_str(); }

int plum_runin_parse(const char* path, plum_runin_summary* out) {
  if (!path || !out) return fail(PLUM_E_ARG, "null argument");
  std::ifstream in(path);
  if (!in) return fail(PLUM_E_ARG, std::string("cannot open ") + path);
  plum::RunIn r;
  std::string err;
  if (!plum::read_sim_block(in, r, err)) return fail(PLUM_E_PARSE, err);
  if (!plum::read_ff_block(in, r, err)) return fail(PLUM_E_PARSE, err);
  std::vector<plum::Molecule> mols;
  int n_bead = 0, n_mol = 0, ids = 0;
  double box[3];
  const std::string dir = dir_of(path);
  if (!plum::read_top(dir + "/" + r.top_file, n_bead, n_mol, box, mols, err)) return fail(PLUM_E_PARSE, err);
  if (!plum::read_crd(dir + "/" + r.crd_file, mols, ids, err)) return fail(PLUM_E_PARSE, err);
  *out = plum_runin_summary{};
  out->n_beads = ids;
  out->n_mols = n_mol;
  std::vector<std::string> syms;
  auto add = [&](const std::string& s) {
    for (auto& t : syms)
      if (t == s) return;
    syms.push_back(s);
  };
  for (auto& t : r.lj_types) add(t.symbol);
  for (auto& t : r.wall_types) add(t.symbol);
  for (auto& m : mols)
    for (auto& b : m.bds) add(b.Symbol());
  out->n_types = (int)syms.size();
  out->npbc = r.npbc;
  out->use_pair = r.use_pair;
  out->use_ewald = r.use_ewald;
  out->use_bond = r.use_bond;
  out->use_ext = r.use_ext;
  out->use_gc = r.use_gc;
  out->dipole_correction = r.dipole_correction;
  for (int k = 0; k < 3; k++) out->box[k] = box[k];
  out->beta = r.beta;
  out->move_size = r.move_size;
  for (int k = 0; k < 5; k++) out->move_prob[k] = r.move_prob[k];
  out->steps = r.steps;
  out->lj_cutoff = r.lj_cutoff;
  out->lB = r.lB;
  out->alpha = r.alpha;
  out->k_bond = r.k_bond;
  out->r0 = r.r0;
  out->rigid_bond = r.rigid_bond;
  out->wall_cut = r.wall_cut;
  out->phantom = r.phantom;
  out->grafted = r.grafted;
  return PLUM_OK;
}

int plum_sim_create(const char* path, unsigned seed, plum_sim** out) {
  if (!path || !out) return fail(PLUM_E_ARG, "null argument");
  *out = nullptr;
  std::ifstream in(path);
  if (!in) return fail(PLUM_E_ARG, std::string("cannot open ") + path);
  plum_sim* s = new plum_sim();
  try {
    s->s = new plum::Simulation(in, seed, dir_of(path));
  } catch (const std::exception& e) {
    delete s;
    return fail(PLUM_E_PARSE, e.what());
  }
  *out = s;
  return PLUM_OK;
}

int plum_sim_run(plum_sim* s, long long steps, long long attempted[5], long long accepted[5]) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  try {
    s->s->Run(steps);
  } catch (const std::exception& e) {
    return fail(PLUM_E_HIP, e.what());
  }
  for (int k = 0; k < 5; k++) {
    if (attempted) attempted[k] = s->s->attempted[k];
    if (accepted) accepted[k] = s->s->accepted[k];
  }
  return PLUM_OK;
}

int plum_sim_run_device(plum_sim* s, long long steps, long long attempted[5], long long accepted[5],
                        double seconds[2]) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  try {
    s->s->RunDevice(steps);
  } catch (const std::exception& e) {
    return fail(PLUM_E_HIP, e.what());
  }
  for (int k = 0; k < 5; k++) {
    if (attempted) attempted[k] = s->s->attempted[k];
    if (accepted) accepted[k] = s->s->accepted[k];
  }
  if (seconds) {
    seconds[0] = s->s->seconds_device;
    seconds[1] = s->s->seconds_gc;
  }
  return PLUM_OK;
}

int plum_sim_gc_stats(plum_sim* s, long long attempted[2], long long accepted[2], int* n_beads, int* n_mols) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  for (int k = 0; k < 2; k++) {
    if (attempted) attempted[k] = s->s->gc_attempted[k];
    if (accepted) accepted[k] = s->s->gc_accepted[k];
  }
  int n = 0;
  for (auto& m : s->s->mols) n += m.Size();
  if (n_beads) *n_beads = n;
  if (n_mols) *n_mols = (int)s->s->mols.size();
  return PLUM_OK;
}

int plum_sim_gc_prescreen(plum_sim* s, int mode, double log_theta) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  if (mode < 0 || mode > 2) return fail(PLUM_E_ARG, "mode must be 0, 1 or 2");
  if (!std::isfinite(log_theta)) return fail(PLUM_E_ARG, "non-finite log_theta");
  s->s->ff.SetGCPrescreen(mode, log_theta);
  return PLUM_OK;
}

int plum_sim_gc_prescreen_stats(plum_sim* s, long long s1_rejected[2], long long s1_passed[2],
                                double* mean_log_rest, double* log_mean_rest, long long* n_rest) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  const auto& p = s->s->ff.PrescreenStats();
  for (int k = 0; k < 2; k++) {
    if (s1_rejected) s1_rejected[k] = p.s1_rejected[k];
    if (s1_passed) s1_passed[k] = p.s1_passed[k];
  }
  if (mean_log_rest) *mean_log_rest = p.n_rest ? p.sum_log_rest / (double)p.n_rest : 0.0;
  if (log_mean_rest) *log_mean_rest = p.n_rest ? p.lse_rest - std::log((double)p.n_rest) : 0.0;
  if (n_rest) *n_rest = p.n_rest;
  return PLUM_OK;
}

int plum_sim_widom(plum_sim* s, double* mean_weight) {
  if (!s || !s->s || !mean_weight) return fail(PLUM_E_ARG, "null argument");
  try {
    *mean_weight = s->s->ff.CalcChemicalPotentialF(s->s->mols, s->s->rng);
  } catch (const std::exception& e) {
    return fail(PLUM_E_HIP, e.what());
  }
  return PLUM_OK;
}

int plum_sim_energy(plum_sim* s, plum_terms* running, plum_terms* full) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  try {
    if (running) *running = s->s->ff.Totals();
    if (full) *full = s->s->ff.FullEnergy();
  } catch (const std::exception& e) {
    return fail(PLUM_E_HIP, e.what());
  }
  return PLUM_OK;
}

int plum_sim_positions(plum_sim* s, double* xyz, int n) {
  if (!s || !s->s || !xyz) return fail(PLUM_E_ARG, "null argument");
  int i = 0;
  for (auto& m : s->s->mols)
    for (auto& b : m.bds) {
      if (i >= n) return fail(PLUM_E_ARG, "buffer too small");
      for (int k = 0; k < 3; k++) xyz[3 * i + k] = b.GetCrd(0, k);
      i++;
    }
  return PLUM_OK;
}

int plum_sim_observables(plum_sim* s, double* rg2, double* ree2) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null simulation");
  if (rg2) *rg2 = s->s->RadiusOfGyration();
  if (ree2) *ree2 = s->s->EndToEndDistance();
  return PLUM_OK;
}

int plum_sim_density_profile(plum_sim* s, double res, double* out, int max_bins, int* n_bins) {
  if (!s || !s->s || !n_bins) return fail(PLUM_E_ARG, "null argument");
  try {
    std::vector<double> v;
    const int nb = s->s->DensityProfile(res, v);
    *n_bins = nb;
    if (out) {
      if (nb > max_bins) return fail(PLUM_E_ARG, "buffer too small");
      std::copy(v.begin(), v.end(), out);
    }
  } catch (const std::exception& e) {
    return fail(PLUM_E_ARG, e.what());
  }
  return PLUM_OK;
}

int plum_sim_write_last(plum_sim* s, const char* prefix) {
  if (!s || !s->s || !prefix) return fail(PLUM_E_ARG, "null argument");
  try {
    s->s->PrintLastCrd(prefix);
    s->s->PrintLastTop(prefix);
  } catch (const std::exception& e) {
    return fail(PLUM_E_ARG, e.what());
  }
  return PLUM_OK;
}

int plum_sim_run_output(plum_sim* s, long long steps, const char* prefix, int device) {
  if (!s || !s->s || !prefix) return fail(PLUM_E_ARG, "null argument");
  if (steps < 0) return fail(PLUM_E_ARG, "steps must be >= 0");
  try {
    s->s->RunWithOutput(steps, prefix, device != 0);
  } catch (const std::exception& e) {
    return fail(PLUM_E_HIP, e.what());
  }
  return PLUM_OK;
}

int plum_sim_pressure_sample(plum_sim* s, plum_vp* sample, double* p_yethiraj, double* p_demiguel,
                             long long* samples) {
  if (!s || !s->s) return fail(PLUM_E_ARG, "null argument");
  try {
    const plum_vp v = s->s->ff.CalcPressureVolScalingHSELSlit(s->s->mols);
    if (sample) *sample = v;
    if (p_yethiraj) *p_yethiraj = s->s->ff.PressureYethiraj();
    if (p_demiguel) *p_demiguel = s->s->ff.PressureDeMiguel();
    if (samples) *samples = s->s->ff.PressureSamples();
  } catch (const std::exception& e) {
    return fail(PLUM_E_HIP, e.what());
  }
  return PLUM_OK;
}

int plum_sim_destroy(plum_sim* s) {
  if (!s) return PLUM_OK;
  delete s->s;
  delete s;
  return PLUM_OK;
}

}  // extern "C"
