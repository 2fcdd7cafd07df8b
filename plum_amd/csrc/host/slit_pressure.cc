// The wall-force slit pressure, ForceField::CalcPressureForceLJELSlit
// (pressure.cc:404-470): the z forces between the phantom wall sites and the
// system (and between the two walls), and the wall potential's force on the
// system's beads.  The reference samples it every sample_freq * 10 steps from
// the settled configuration; it runs here on the host from the coordinates the
// host already holds for its other observables.
//
// The pair forces are the reference's (PairForce, PairForceZReal,
// PairForceZRepl, BeadForceOnWall) with one change of order: the reciprocal
// z force sum_j sum_k kz e(k) sin(k.(r_i - r_j)) of a wall site i is formed
// as sum_k kz e(k) Im(e^{i k.r_i} S_c(k)*) from the structure factor S_c of
// each class c of partner (ions, chain beads, the other wall's sites), so the
// cost is O((sites + beads) N_k) instead of O(sites beads N_k).  The k grid,
// cutoffs and images are the engine's (plum_info).
#include <cmath>
#include <complex>
#include <vector>

#include "plum_host.h"

namespace {

constexpr double kVeryLarge = 1e8;          // kVeryLargeEnergy (misc.h)
constexpr double kPiRef = 3.14159265359;    // kPi as the reference spells it
constexpr double k213 = 1.25992104989;
constexpr double k216 = 1.12246204831;
constexpr double kWallA = 2.59807621135;

// Bead::GetDistVector (bead.cc): d = b - a, minimum image in the periodic
// dimensions.
void dist_vector(const double* a, const double* b, const double box[3], int npbc, double d[3]) {
  for (int k = 0; k < 3; k++) {
    d[k] = b[k] - a[k];
    if (k < npbc) d[k] -= box[k] * std::round(d[k] / box[k]);
  }
}

// Bead::BBDist (bead.cc:162-176).
double bb_dist(const double* a, const double* b, const double box[3], int npbc) {
  double s = 0;
  for (int k = 0; k < 3; k++) {
    double d = a[k] - b[k];
    if (k < npbc) {
      d -= box[k] * std::round(d / box[k]);
      if (std::fabs(d) > 0.5 * box[k]) d = box[k] - std::fabs(d);
    }
    s += d * d;
  }
  return std::sqrt(s);
}

// PotentialTruncatedLJ::PairForce (potential_truncated_lj.cc:87-121): the
// magnitude along the pair; beyond a full-LJ cutoff the reference's
// expression with (sigma / lj_cutoff)^6 is kept.
double lj_force(const plum_params& p, double r, int ta, int tb) {
  const double sigma = (p.lj_sigma[ta] + p.lj_sigma[tb]) / 2;
  const double eps = std::sqrt(p.lj_eps[ta] * p.lj_eps[tb]);
  if (r <= 0) return kVeryLarge;
  if (p.lj_cutoff < 0) {
    if (r < k216 * sigma) {
      const double r6 = std::pow(sigma / r, 6);
      return 4 * eps * (r6 * r6 * 12 / r - r6 * 6 / r);
    }
    return 0.0;
  }
  const double r6 = std::pow(sigma / (r < p.lj_cutoff ? r : p.lj_cutoff), 6);
  return 4 * eps * (r6 * r6 * 12 / r - r6 * 6 / r);
}

// PotentialEwaldCoul::PairForceZReal (potential_ewald_coul.cc:261-293): the z
// force on wall site a from bead b over the real-space images.
double real_force_z(const plum_params& p, const plum_info& e, const double* a, double qa, const double* b,
                    double qb) {
  if (qa * qb == 0) return 0.0;
  double r[3];
  dist_vector(b, a, e.box_e, p.npbc, r);
  const double pref = p.lB * qa * qb, pref2 = 2 * std::sqrt(p.alpha / kPiRef), sa = std::sqrt(p.alpha);
  double f = 0;
  for (int i = -e.real_cell[0]; i <= e.real_cell[0]; i++)
    for (int j = -e.real_cell[1]; j <= e.real_cell[1]; j++)
      for (int k = -e.real_cell[2]; k <= e.real_cell[2]; k++) {
        const double x = r[0] + i * e.box_e[0], y = r[1] + j * e.box_e[1], z = r[2] + k * e.box_e[2];
        const double d = std::sqrt(x * x + y * y + z * z);
        if (d > 0 && d <= e.real_cutoff)
          f += pref * ((pref2 * std::exp(-p.alpha * d * d)) + (std::erfc(sa * d) / d)) * z / (d * d);
      }
  return f;
}

// PotentialTruncatedLJWall::BeadForceOnWall (potential_truncated_lj_wall.cc:
// 136-214), branch for branch (the FENE term and the full-LJ branch's last
// term as the reference writes them).
double wall_force(const plum_params& p, double z, int t) {
  const double Lz = p.box[2];
  const double sigma = p.wall_sigma[t], eps = p.wall_eps[t];
  const int kind = p.type_kind[t];
  const double R0 = 3 * k213 * sigma, K = 1, m_cut = p.wall_cut;
  auto term = [&](double s, double d) {
    const double r3 = std::pow(s / d, 3);
    return r3;
  };
  double f = 0;
  if (eps == 0) return 0.0;
  if (z <= 0 || z >= Lz) return kVeryLarge;
  if (m_cut < 0 && kind != 2 && kind != 1) {
    if (z < k213 * sigma) {
      const double r3 = term(sigma, z);
      f += kWallA * eps * (r3 * r3 * 6 / z - r3 * 3 / z);
    }
    if (Lz - z < k213 * sigma) {
      const double r3 = term(sigma, Lz - z);
      f += kWallA * eps * (r3 * r3 * 6 / (Lz - z) - r3 * 3 / (Lz - z));
    }
  } else if (kind == 1) {
    const double r3 = term(sigma, z);
    f += kWallA * eps * (r3 * r3 * 6 / z - r3 * 3 / z);
    f += -K * R0 * z / (1 - std::pow(z / R0, 2));
    if (Lz - z < k213 * sigma) {
      const double r3b = term(sigma, Lz - z);
      f += kWallA * eps * (r3b * r3b * 6 / (Lz - z) - r3b * 3 / (Lz - z));
    }
  } else if (kind == 2) {
    const double r3 = term(sigma, Lz - z);
    f += kWallA * eps * (r3 * r3 * 6 / (Lz - z) - r3 * 3 / (Lz - z));
    f += -K * R0 * (Lz - z) / (1 - std::pow((Lz - z) / R0, 2));
    if (z < k213 * sigma) {
      const double r3b = term(sigma, z);
      f += kWallA * eps * (r3b * r3b * 6 / z - r3b * 3 / z);
    }
  } else {
    const double r3l = term(sigma, z < m_cut ? z : m_cut);
    f += kWallA * eps * (r3l * r3l * 6 / z - r3l * 3 / z);
    if (Lz - z < m_cut) {
      const double r3 = term(sigma, Lz - z);
      f += kWallA * eps * (r3 * r3 * 6 / (Lz - z) - r3 * 3 / (Lz - z));
    } else {
      const double r3 = term(sigma, m_cut);
      f += kWallA * eps * (r3 * r3 * 6 / z - r3 * 3 / z);
    }
  }
  return f;
}

}  // namespace

extern "C" int plum_slit_forces(const plum_params* pp, const plum_info* ee, int n_beads, const double* xyz,
                                const double* q, const int* type, int n_mols, const int* mol_start, int phantom,
                                double f[6]) {
  if (!pp || !ee || !xyz || !q || !type || !mol_start || !f || n_beads < 0 || n_mols < 0 || phantom < 0 ||
      phantom > n_mols)
    return PLUM_E_ARG;
  const plum_params& p = *pp;
  const plum_info& e = *ee;
  if (p.n_types < 1 || p.n_types > PLUM_MAX_TYPES || mol_start[0] != 0 || mol_start[n_mols] != n_beads)
    return PLUM_E_ARG;
  for (int m = 0; m < n_mols; m++)
    if (mol_start[m + 1] <= mol_start[m]) return PLUM_E_ARG;
  for (int i = 0; i < n_beads; i++)
    if (type[i] < 0 || type[i] >= p.n_types) return PLUM_E_ARG;
  for (int k = 0; k < 6; k++) f[k] = 0;
  const int half = phantom / 2;
  // partner class of molecule j (pressure.cc:430-434): 0 ion, 1 chain, 2 the
  // other wall
  auto cls = [&](int j) { return j >= phantom ? (mol_start[j + 1] - mol_start[j] > 1 ? 1 : 0) : 2; };
  // LJ and real-space Ewald, pair by pair (pressure.cc:418-457)
  for (int i = 0; i < phantom; i++) {
    const int wi = mol_start[i];
    const double* pw = xyz + 3 * wi;
    for (int j = half; j < n_mols; j++) {
      if (!(i < half || j >= phantom)) continue;
      const double C = (i < half && j >= phantom) ? -0.5 : (j >= phantom ? 0.5 : -1.0);
      const int c = cls(j);
      for (int k = mol_start[j]; k < mol_start[j + 1]; k++) {
        const double* pb = xyz + 3 * k;
        if (p.use_pair) {
          double r[3];
          dist_vector(pb, pw, p.box, p.npbc, r);
          const double rr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
          f[c] += C * (r[2] / rr) * lj_force(p, bb_dist(pb, pw, p.box, p.npbc), type[k], type[wi]);
        }
        if (p.use_ewald) f[3 + c] += C * real_force_z(p, e, pw, q[wi], pb, q[k]);
      }
    }
  }
  // reciprocal-space Ewald through the structure factors of the partner
  // classes (PairForceZRepl, potential_ewald_coul.cc:297-324)
  if (p.use_ewald && phantom > 0) {
    const int cx = 2 * e.repl_cell[0] + 1, cy = 2 * e.repl_cell[1] + 1, cz = 2 * e.repl_cell[2] + 1;
    std::vector<double> kv[3];
    const int cnt[3] = {cx, cy, cz};
    for (int d = 0; d < 3; d++) {
      kv[d].resize(cnt[d]);
      for (int l = -e.repl_cell[d]; l <= e.repl_cell[d]; l++)
        kv[d][l + e.repl_cell[d]] = (double)(l * 2) * kPiRef / e.box_e[d];
    }
    struct KPoint {
      int ix, iy, iz;
      double wz;   // kz e(k) with e(k) = exp(-k^2 / 4 alpha) / k^2
    };
    std::vector<KPoint> ks;
    for (int ix = 0; ix < cx; ix++)
      for (int iy = 0; iy < cy; iy++)
        for (int iz = 0; iz < cz; iz++) {
          const double k2 = kv[0][ix] * kv[0][ix] + kv[1][iy] * kv[1][iy] + kv[2][iz] * kv[2][iz];
          if (k2 > 0 && k2 <= e.repl_cutoff)
            ks.push_back({ix, iy, iz, kv[2][iz] * std::exp(-k2 / (4 * p.alpha)) / k2});
        }
    // per-dimension phases e^{-i k_d x_d} of a position (all grid values)
    typedef std::complex<double> cd;
    auto phases = [&](const double* x, std::vector<cd>* ph) {
      for (int d = 0; d < 3; d++) {
        ph[d].resize(cnt[d]);
        for (int l = 0; l < cnt[d]; l++) ph[d][l] = std::polar(1.0, -kv[d][l] * x[d]);
      }
    };
    // S_c(k) = sum over class c of q e^{-i k.r}
    std::vector<cd> S[3];
    for (int c = 0; c < 3; c++) S[c].assign(ks.size(), cd(0, 0));
    std::vector<cd> ph[3];
    for (int j = half; j < n_mols; j++) {
      const int c = cls(j);
      for (int k = mol_start[j]; k < mol_start[j + 1]; k++) {
        if (q[k] == 0) continue;
        phases(xyz + 3 * k, ph);
        for (size_t m = 0; m < ks.size(); m++)
          S[c][m] += q[k] * ph[0][ks[m].ix] * ph[1][ks[m].iy] * ph[2][ks[m].iz];
      }
    }
    const double pref = p.lB * 4 * kPiRef / e.vol_e;
    for (int i = 0; i < phantom; i++) {
      const int wi = mol_start[i];
      if (q[wi] == 0) continue;
      phases(xyz + 3 * wi, ph);
      // sum_k kz e(k) sin(k.(r_i - r_j)) q_j = sum_k kz e(k) Im(conj(phase_i) S_c)
      double acc[3] = {0, 0, 0};
      for (size_t m = 0; m < ks.size(); m++) {
        const cd ei = std::conj(ph[0][ks[m].ix] * ph[1][ks[m].iy] * ph[2][ks[m].iz]);
        for (int c = 0; c < 3; c++) acc[c] += ks[m].wz * (ei * S[c][m]).imag();
      }
      for (int c = 0; c < 3; c++) {
        if (c == 2 && i >= half) continue;   // the wall-wall pairs once, from the z = 0 wall
        const double C = c == 2 ? -1.0 : (i < half ? -0.5 : 0.5);
        f[3 + c] += C * pref * q[wi] * acc[c];
      }
    }
  }
  // the wall potential on the system's beads (pressure.cc:459-470)
  if (p.use_ext)
    for (int i = phantom; i < n_mols; i++) {
      const int c = mol_start[i + 1] - mol_start[i] > 1 ? 1 : 0;
      for (int j = mol_start[i]; j < mol_start[i + 1]; j++) f[c] += 0.5 * wall_force(p, xyz[3 * j + 2], type[j]);
    }
  return PLUM_OK;
}
