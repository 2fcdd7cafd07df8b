// Seeded synthetic systems in the reference's data model (chains as random
// walks, ions uniform), for workloads that have no input deck.  No reference
// counterpart: the reference reads its configuration from a crd file
// (simulation.cc:582-646).
#include <cmath>
#include <cstdint>
#include <random>
#include <unordered_map>
#include <vector>

#include "plum_host.h"

namespace {

struct HashGrid {
  double box[3], h[3];
  int n[3];
  bool periodic[3];
  std::unordered_map<long long, std::vector<int>> cells;
  const std::vector<double>* xyz;

  long long key(int cx, int cy, int cz) const { return ((long long)cx * n[1] + cy) * n[2] + cz; }

  void coords(const double* p, int c[3]) const {
    for (int d = 0; d < 3; d++) {
      double w = p[d];
      if (periodic[d]) w -= box[d] * std::floor(w / box[d]);
      int k = (int)std::floor(w / h[d]);
      c[d] = k < 0 ? 0 : (k >= n[d] ? n[d] - 1 : k);
    }
  }

  void add(int i) {
    int c[3];
    coords(&(*xyz)[3 * (size_t)i], c);
    cells[key(c[0], c[1], c[2])].push_back(i);
  }

  bool too_close(const double* p, double r) const {
    int c[3];
    coords(p, c);
    const double r2 = r * r;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          int k[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
          bool ok = true;
          for (int d = 0; d < 3; d++) {
            if (periodic[d]) k[d] = (k[d] % n[d] + n[d]) % n[d];
            else if (k[d] < 0 || k[d] >= n[d]) ok = false;
          }
          if (!ok) continue;
          auto it = cells.find(key(k[0], k[1], k[2]));
          if (it == cells.end()) continue;
          for (int j : it->second) {
            double s = 0;
            for (int d = 0; d < 3; d++) {
              double dd = p[d] - (*xyz)[3 * (size_t)j + d];
              if (periodic[d]) dd -= box[d] * std::round(dd / box[d]);
              s += dd * dd;
            }
            if (s < r2) return true;
          }
        }
    return false;
  }
};

// Marsaglia (1972) point on the unit sphere.
void sphere(std::mt19937_64& g, double v[3]) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (;;) {
    const double r1 = 1 - 2 * u(g), r2 = 1 - 2 * u(g);
    const double s = r1 * r1 + r2 * r2;
    if (s <= 1) {
      const double h = 2 * std::sqrt(1 - s);
      v[0] = r1 * h;
      v[1] = r2 * h;
      v[2] = 1 - 2 * s;
      return;
    }
  }
}

}  // namespace

extern "C" {

int plum_synth_count(const plum_synth_spec* s, int* n_beads, int* n_mols) {
  if (!s || !n_beads || !n_mols) return PLUM_E_ARG;
  if (s->n_chains < 0 || s->chain_len < 1 || s->n_counterions < 0 || s->n_salt_pairs < 0) return PLUM_E_ARG;
  const long long nb = (long long)s->n_chains * s->chain_len + s->n_counterions + 2LL * s->n_salt_pairs;
  const long long nm = (long long)s->n_chains + s->n_counterions + 2LL * s->n_salt_pairs;
  if (nb > (1LL << 30)) return PLUM_E_ARG;
  *n_beads = (int)nb;
  *n_mols = (int)nm;
  return PLUM_OK;
}

int plum_synth_system(const plum_synth_spec* s, double* xyz_out, double* q, int* type, int* mol_start) {
  int nb = 0, nm = 0;
  if (plum_synth_count(s, &nb, &nm) != PLUM_OK || !xyz_out || !q || !type || !mol_start) return PLUM_E_ARG;
  for (int d = 0; d < 3; d++)
    if (!(s->box[d] > 0)) return PLUM_E_ARG;
  std::mt19937_64 g(s->seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<double> xyz;
  xyz.reserve(3 * (size_t)nb);
  HashGrid grid;
  const double cell = std::max(s->min_sep, 1.0);
  for (int d = 0; d < 3; d++) {
    grid.box[d] = s->box[d];
    grid.n[d] = std::max(1, (int)(s->box[d] / cell));
    grid.h[d] = s->box[d] / grid.n[d];
    grid.periodic[d] = d < 2 || s->npbc == 3;
  }
  grid.xyz = &xyz;
  const bool zr = s->z_hi > s->z_lo;
  const double zlo = zr ? s->z_lo : 0.0, zhi = zr ? s->z_hi : s->box[2];
  int n = 0, m = 0;
  mol_start[0] = 0;
  auto push = [&](const double* p, double qq, int t) {
    xyz.push_back(p[0]);
    xyz.push_back(p[1]);
    xyz.push_back(p[2]);
    q[n] = qq;
    type[n] = t;
    grid.add(n);
    n++;
  };
  // a chain that walks into a dead end is regrown from a new start
  std::vector<double> ch;
  for (int c = 0; c < s->n_chains; c++) {
    bool grown = false;
    for (int attempt = 0; attempt < 1000 && !grown; attempt++) {
      ch.clear();
      double p[3];
      int tries = 0;
      for (;;) {
        p[0] = u(g) * s->box[0];
        p[1] = u(g) * s->box[1];
        p[2] = s->graft_z >= 0 ? s->graft_z : zlo + u(g) * (zhi - zlo);
        if (!grid.too_close(p, s->min_sep)) break;
        if (++tries > 100000) return PLUM_E_CAPACITY;
      }
      ch.insert(ch.end(), p, p + 3);
      bool stuck = false;
      for (int k = 1; k < s->chain_len && !stuck; k++) {
        tries = 0;
        for (;;) {
          double v[3], t[3];
          sphere(g, v);
          const double* last = &ch[ch.size() - 3];
          for (int d = 0; d < 3; d++) t[d] = last[d] + s->r0 * v[d];
          bool ok = !zr || (t[2] > zlo && t[2] < zhi);
          ok = ok && !grid.too_close(t, s->min_sep);
          for (size_t b = 0; ok && b + 3 <= ch.size(); b += 3) {
            double d2 = 0;
            for (int d = 0; d < 3; d++) {
              double dd = t[d] - ch[b + d];
              if (grid.periodic[d]) dd -= s->box[d] * std::round(dd / s->box[d]);
              d2 += dd * dd;
            }
            ok = d2 >= s->min_sep * s->min_sep;
          }
          if (ok) {
            ch.insert(ch.end(), t, t + 3);
            break;
          }
          if (++tries > 2000) {
            stuck = true;
            break;
          }
        }
      }
      if (stuck) continue;
      for (int k = 0; k < s->chain_len; k++) push(&ch[3 * (size_t)k], s->bead_q, (k == 0 && s->graft_z >= 0) ? 2 : 0);
      grown = true;
    }
    if (!grown) return PLUM_E_CAPACITY;
    mol_start[++m] = n;
  }
  for (int i = 0; i < s->n_counterions + 2 * s->n_salt_pairs; i++) {
    const double qq = i < s->n_counterions ? s->counterion_q : ((i - s->n_counterions) % 2 == 0 ? 1.0 : -1.0);
    double p[3];
    int tries = 0;
    for (;;) {
      p[0] = u(g) * s->box[0];
      p[1] = u(g) * s->box[1];
      p[2] = zlo + u(g) * (zhi - zlo);
      if (!grid.too_close(p, s->min_sep)) break;
      if (++tries > 100000) return PLUM_E_CAPACITY;
    }
    push(p, qq, 1);
    mol_start[++m] = n;
  }
  for (size_t i = 0; i < xyz.size(); i++) xyz_out[i] = xyz[i];
  return PLUM_OK;
}

}  // extern "C"
