/*
 * plum_oracle.c -- FP64 CPU restatement of the reference trial-move energy
 * path.  TEST INFRASTRUCTURE ONLY: loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product.  It follows the reference algorithm literally (pairwise
 * multi-image real space, pairwise full-grid k-space, per-pair old/new
 * differences, the same early exits and the dipole lag), so that the HIP
 * engine, which regroups the same sums (structure factor, cell lists), is
 * checked against the reference's own arithmetic.
 *
 * Pinning of each part (the checks that would catch a dropped term, a wrong
 * sign/index or a transposed operand), all in tests/test_oracle.py:
 *   cutoff derivation ........ the four reference example logs (r_c, k_c^2,
 *                               rounded cell counts) -> tests/golden/
 *   LJ / WCA ................. closed forms at chosen r, continuity at the
 *                               cutoff, sigma/epsilon mixing
 *   real space ............... brute-force image sum on tiny boxes
 *   k space .................. pairwise == |S(k)|^2 identity; Madelung
 *                               constants (NaCl, CsCl) from the full Ewald
 *                               total, alpha-independence
 *   self, dipole, spring ..... closed forms
 *   6-3 wall + FENE .......... closed forms incl. the L/R branches
 *   dE bookkeeping ........... dE == E_scratch(after) - E_scratch(before)
 *                               for every term except the dipole lag, and the
 *                               lag itself against its two-step definition
 *   mt19937 .................. C++ [rand.predef] known answer (10000th
 *                               output of the default-seeded engine)
 *   move generators .......... invariants (bond lengths, rigid rotation,
 *                               translation); the crankshaft rotation is a
 *                               Rodrigues restatement of Eigen's AngleAxis
 *                               (Eigen is absent) -> bit-level parity of
 *                               proposals is unpinned, geometry is pinned.
 */
#include "plum_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* utilities/constants.h:11-62 */
static const double kVeryLargeEnergy = 1E+8;
static const double kPi = 3.14159265359;
static const double kEwaldCutoff = 1E-7;
static const double k213 = 1.25992104989;
static const double k216 = 1.12246204831;
static const int kDiCorrection = 3;

struct orc_sys {
  orc_params p;
  orc_ewald_info e;
  int n, n_mols;
  double *cur, *trial;     /* 3n interleaved */
  double *q;
  int *type, *mol_of, *mol_start;
  unsigned char *moved;
  /* k tables (potential_ewald_coul.cc:78-115) */
  int ceto[3];
  double *kx, *ky, *kz, *k2, *ek2;
  /* running totals and pending differences (force_field.cc:436-451) */
  double E_lj, E_real, E_kspace, E_self, E_dipole, E_bond, E_wall;
  double cur_dipl, trial_dipl;
  double pend[ORC_NTERMS];
  int pend_valid;
  double *bond_cur;        /* per-molecule spring energy cache */
};

/* ------------------------------------------------------------------ */
/* geometry (misc.cc:43-56, bead.cc:162-176)                           */
/* ------------------------------------------------------------------ */

/* GetDistVector: vector from b1 to b2 on TRIAL coordinates, minimum image
   with round() in the first npbc dimensions (misc.cc:43-56). */
static void get_dist_vector(const double* b1, const double* b2,
                            const double box[3], int npbc, double d[3]) {
  for (int i = 0; i < 3; i++) {
    double di = b2[i] - b1[i];
    if (i < npbc) di -= box[i] * round(di / box[i]);
    d[i] = di;
  }
}

/* Bead::BBDist on trial coordinates (bead.cc:162-176). */
static double bb_dist(const double* a, const double* b, const double box[3],
                      int npbc) {
  double dist = 0;
  for (int i = 0; i < 3; i++) {
    double di = a[i] - b[i];
    if (i < npbc) {
      di -= box[i] * round(di / box[i]);
      if (fabs(di) > 0.5 * box[i]) di = box[i] - fabs(di);
    }
    dist += di * di;
  }
  return sqrt(dist);
}

/* ------------------------------------------------------------------ */
/* Ewald set-up (potential_ewald_coul.cc:29-115)                        */
/* ------------------------------------------------------------------ */
void orc_derive_cutoffs(const double box_in[3], double lB, double alpha,
                        int dipole, orc_ewald_info* out) {
  double box[3] = {box_in[0], box_in[1], box_in[2]};
  /* potential_ewald_coul.cc:43-49 -- slab padding */
  if (dipole) {
    double min_padding = 150;
    if (box[2] * kDiCorrection > min_padding) box[2] = box[2] * kDiCorrection;
    else box[2] += min_padding;
  }
  for (int i = 0; i < 3; i++) out->box_e[i] = box[i];
  out->vol_e = box[0] * box[1] * box[2];
  /* :56-61 real-space cutoff, one unit length at a time */
  double real_cutoff = 1;
  while (0.5 * lB * 1 * 1 * erfc(sqrt(alpha) * real_cutoff) / real_cutoff >
         kEwaldCutoff)
    real_cutoff += 1;
  out->real_cutoff = real_cutoff;
  for (int i = 0; i < 3; i++) out->real_cell[i] = (int)ceil(real_cutoff / box[i]);
  /* :66-75 reciprocal cutoff, alpha at a time (accumulated by addition) */
  double repl_cutoff = alpha;
  double min_box_vol = pow(box[0], 3);
  while (lB * 1 * 1 / (2 * kPi * min_box_vol) * (4 * kPi * kPi) *
             exp(-repl_cutoff / (4 * alpha)) / repl_cutoff >
         kEwaldCutoff)
    repl_cutoff += alpha;
  out->repl_cutoff = repl_cutoff;
  for (int i = 0; i < 3; i++)
    out->repl_cell[i] = (int)ceil(sqrt(repl_cutoff) * box[i] / (2 * kPi));
  out->n_k_full = -1;
}

static void build_k_tables(orc_sys* s) {
  const double* box = s->e.box_e;
  const int* rc = s->e.repl_cell;
  for (int d = 0; d < 3; d++) s->ceto[d] = 2 * rc[d] + 1;
  int n3 = s->ceto[0] * s->ceto[1] * s->ceto[2];
  s->kx = malloc(sizeof(double) * s->ceto[0]);
  s->ky = malloc(sizeof(double) * s->ceto[1]);
  s->kz = malloc(sizeof(double) * s->ceto[2]);
  s->k2 = malloc(sizeof(double) * n3);
  s->ek2 = malloc(sizeof(double) * n3);
  /* potential_ewald_coul.cc:89-96: l * 2*kPi / L with an int product first */
  for (int l = -rc[0]; l <= rc[0]; l++) s->kx[l + rc[0]] = (double)(l * 2) * kPi / box[0];
  for (int l = -rc[1]; l <= rc[1]; l++) s->ky[l + rc[1]] = (double)(l * 2) * kPi / box[1];
  for (int l = -rc[2]; l <= rc[2]; l++) s->kz[l + rc[2]] = (double)(l * 2) * kPi / box[2];
  int nk = 0;
  /* :102-115 */
  for (int ix = 0; ix < s->ceto[0]; ix++)
    for (int iy = 0; iy < s->ceto[1]; iy++)
      for (int iz = 0; iz < s->ceto[2]; iz++) {
        int idx = s->ceto[1] * s->ceto[2] * ix + s->ceto[2] * iy + iz;
        s->k2[idx] = s->kx[ix] * s->kx[ix] + s->ky[iy] * s->ky[iy] + s->kz[iz] * s->kz[iz];
        s->ek2[idx] = exp(-s->k2[idx] / (4 * s->p.alpha)) / s->k2[idx];
        if (s->k2[idx] > 0 && s->k2[idx] <= s->e.repl_cutoff) nk++;
      }
  s->e.n_k_full = nk;
}

/* ------------------------------------------------------------------ */
/* pair / bead energies                                                */
/* ------------------------------------------------------------------ */

/* PotentialTruncatedLJ::PairEnergy (potential_truncated_lj.cc:49-85) */
static double lj_pair_xyz(const orc_sys* s, const double* a, int ta,
                          const double* b, int tb) {
  double energy = 0;
  double r = bb_dist(a, b, s->p.box, s->p.npbc);
  double sigma = (s->p.lj_sigma[ta] + s->p.lj_sigma[tb]) / 2;
  double epsilon = sqrt(s->p.lj_eps[ta] * s->p.lj_eps[tb]);
  double r6 = 0;
  if (r <= 0) r6 = -1;
  if (r6 == -1) {
    energy = kVeryLargeEnergy;
  } else if (s->p.lj_cutoff < 0) {
    double r6_ref = pow(1.0 / k216, 6);
    double energy_ref = 4 * epsilon * (r6_ref * r6_ref - r6_ref);
    if (r < k216 * sigma) {
      r6 = pow(sigma / r, 6);
      energy = 4 * epsilon * (r6 * r6 - r6) - energy_ref;
    }
  } else if (r < s->p.lj_cutoff) {
    double r6_ref = pow(sigma / s->p.lj_cutoff, 6);
    double energy_ref = 4 * epsilon * (r6_ref * r6_ref - r6_ref);
    r6 = pow(sigma / r, 6);
    energy = 4 * epsilon * (r6 * r6 - r6) - energy_ref;
  }
  return energy;
}

/* PotentialEwaldCoul::PairEnergyReal (potential_ewald_coul.cc:134-164) */
static double real_pair_xyz(const orc_sys* s, const double* a, double qa,
                            const double* b, double qb) {
  double energy = 0;
  if (qa * qb == 0) return energy;
  double dist[3];
  get_dist_vector(a, b, s->e.box_e, s->p.npbc, dist);
  double prefactor = s->p.lB * qa * qb;
  const int* c = s->e.real_cell;
  const double* L = s->e.box_e;
  double sa = sqrt(s->p.alpha);
  for (int i = -c[0]; i <= c[0]; i++)
    for (int j = -c[1]; j <= c[1]; j++)
      for (int k = -c[2]; k <= c[2]; k++) {
        double rv0 = dist[0] + i * L[0];
        double rv1 = dist[1] + j * L[1];
        double rv2 = dist[2] + k * L[2];
        double r = sqrt(rv0 * rv0 + rv1 * rv1 + rv2 * rv2);
        if (r > 0 && r <= s->e.real_cutoff) energy += prefactor * erfc(sa * r) / r;
      }
  return energy;
}

/* PotentialEwaldCoul::PairEnergyRepl (potential_ewald_coul.cc:198-225) */
static double repl_pair_xyz(const orc_sys* s, const double* a, double qa,
                            const double* b, double qb) {
  double energy = 0;
  if (qa * qb == 0) return energy;
  double r[3];
  get_dist_vector(a, b, s->e.box_e, s->p.npbc, r);
  double prefactor = s->p.lB * qa * qb / (kPi * s->e.vol_e) * (4 * kPi * kPi);
  for (int lx = 0; lx < s->ceto[0]; lx++)
    for (int ly = 0; ly < s->ceto[1]; ly++)
      for (int lz = 0; lz < s->ceto[2]; lz++) {
        int idx = s->ceto[1] * s->ceto[2] * lx + s->ceto[2] * ly + lz;
        if (s->k2[idx] > 0 && s->k2[idx] <= s->e.repl_cutoff)
          energy += prefactor * s->ek2[idx] *
                    cos(s->kx[lx] * r[0] + s->ky[ly] * r[1] + s->kz[lz] * r[2]);
      }
  return energy;
}

/* PotentialEwaldCoul::SelfEnergy (potential_ewald_coul.cc:253-257) */
static double self_energy(const orc_sys* s, double q) {
  return -s->p.lB * sqrt(s->p.alpha / kPi) * q * q;
}

/* PotentialEwaldCoul::DipoleE (potential_ewald_coul.cc:415-428): CURRENT z. */
static double dipole_energy_current(const orc_sys* s) {
  double Mz = 0;
  for (int i = 0; i < s->n; i++) Mz += s->q[i] * s->cur[3 * i + 2];
  return s->p.lB * 2 * kPi / (s->e.vol_e / 1.0) * Mz * Mz;
}

/* PotentialSpring::MoleculeEnergy on trial coordinates, no PBC
   (potential_spring.cc:22-47). */
static double spring_molecule(const orc_sys* s, int m) {
  double e = 0;
  static const double nobox[3] = {0, 0, 0};
  for (int i = s->mol_start[m]; i < s->mol_start[m + 1] - 1; i++) {
    double r = bb_dist(&s->trial[3 * i], &s->trial[3 * (i + 1)], nobox, 0);
    e += 0.5 * s->p.k_bond * (r - s->p.r0) * (r - s->p.r0);
  }
  return e;
}

/* PotentialTruncatedLJWall::BeadEnergy on trial z
   (potential_truncated_lj_wall.cc:46-134). box is the unpadded box. */
static double wall_bead_z(const orc_sys* s, double z, int t) {
  double energy = 0;
  const double Lz = s->p.box[2];
  double sigma = s->p.wall_sigma[t];
  double epsilon = s->p.wall_eps[t];
  int kind = s->p.type_kind[t];
  double R0 = 3 * k213 * sigma;
  double K = 1;
  double m_cut = s->p.wall_cut;
  if (epsilon == 0) {
    energy = 0;
  } else if (z <= 0 || z >= Lz) {
    return kVeryLargeEnergy;
  } else {
    if (m_cut < 0 && kind != ORC_KIND_R && kind != ORC_KIND_L) {
      double r3_ref = pow(1.0 / k213, 3);
      double energy_ref = 2.59807621135 * epsilon * (r3_ref * r3_ref - r3_ref);
      if (z < k213 * sigma) {
        double r3 = pow(sigma / z, 3);
        energy += 2.59807621135 * epsilon * (r3 * r3 - r3) - energy_ref;
      }
      if (Lz - z < k213 * sigma) {
        double r3 = pow(sigma / (Lz - z), 3);
        energy += 2.59807621135 * epsilon * (r3 * r3 - r3) - energy_ref;
      }
    } else if (kind == ORC_KIND_L) {
      double r3 = pow(sigma / z, 3);
      energy += 2.59807621135 * epsilon * (r3 * r3 - r3);
      energy += -0.5 * K * R0 * R0 * log(1 - pow(z / R0, 2));
      if (Lz - z < k213 * sigma) {
        double r3b = pow(sigma / (Lz - z), 3);
        energy += 2.59807621135 * epsilon * (r3b * r3b - r3b);
      } else {
        double r3b = pow(1.0 / k213, 3);
        energy += 2.59807621135 * epsilon * (r3b * r3b - r3b);
      }
    } else if (kind == ORC_KIND_R) {
      double r3 = pow(sigma / (Lz - z), 3);
      energy += 2.59807621135 * epsilon * (r3 * r3 - r3);
      energy += -0.5 * K * R0 * R0 * log(1 - pow((Lz - z) / R0, 2));
      if (z < k213 * sigma) {
        double r3b = pow(sigma / z, 3);
        energy += 2.59807621135 * epsilon * (r3b * r3b - r3b);
      } else {
        double r3b = pow(1.0 / k213, 3);
        energy += 2.59807621135 * epsilon * (r3b * r3b - r3b);
      }
    } else {
      double r3_ref = pow(1.0 / m_cut, 3);
      double energy_ref = 2.59807621135 * epsilon * (r3_ref * r3_ref - r3_ref);
      if (z < m_cut) {
        double r3 = pow(sigma / z, 3);
        energy += 2.59807621135 * epsilon * (r3 * r3 - r3) - energy_ref;
      }
      if (Lz - z < m_cut) {
        double r3 = pow(sigma / (Lz - z), 3);
        energy += 2.59807621135 * epsilon * (r3 * r3 - r3) - energy_ref;
      }
    }
  }
  return energy;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */
orc_sys* orc_create(const orc_params* p, int n, const double* xyz,
                    const double* q, const int* type, int n_mols,
                    const int* mol_start) {
  return orc_create_ex(p, n, xyz, q, type, n_mols, mol_start, 1);
}

orc_sys* orc_create_ex(const orc_params* p, int n, const double* xyz,
                       const double* q, const int* type, int n_mols,
                       const int* mol_start, int init_energy) {
  orc_sys* s = calloc(1, sizeof(orc_sys));
  s->p = *p;
  s->n = n;
  s->n_mols = n_mols;
  s->cur = malloc(sizeof(double) * 3 * (n ? n : 1));
  s->trial = malloc(sizeof(double) * 3 * (n ? n : 1));
  memcpy(s->cur, xyz, sizeof(double) * 3 * n);
  memcpy(s->trial, xyz, sizeof(double) * 3 * n);
  s->q = malloc(sizeof(double) * (n ? n : 1));
  memcpy(s->q, q, sizeof(double) * n);
  s->type = malloc(sizeof(int) * (n ? n : 1));
  memcpy(s->type, type, sizeof(int) * n);
  s->mol_start = malloc(sizeof(int) * (n_mols + 1));
  memcpy(s->mol_start, mol_start, sizeof(int) * (n_mols + 1));
  s->mol_of = malloc(sizeof(int) * (n ? n : 1));
  for (int m = 0; m < n_mols; m++)
    for (int i = mol_start[m]; i < mol_start[m + 1]; i++) s->mol_of[i] = m;
  s->moved = calloc(n ? n : 1, 1);
  s->bond_cur = calloc(n_mols ? n_mols : 1, sizeof(double));
  if (p->use_ewald) {
    orc_derive_cutoffs(p->box, p->lB, p->alpha, p->dipole, &s->e);
    build_k_tables(s);
  } else {
    for (int i = 0; i < 3; i++) s->e.box_e[i] = p->box[i];
    s->e.vol_e = p->box[0] * p->box[1] * p->box[2];
  }
  /* ForceField::InitializeEnergy -> running totals (force_field.cc:365-381) */
  double t[ORC_NTERMS] = {0, 0, 0, 0, 0, 0, 0};
  if (init_energy) orc_energy_full(s, t);
  /* the dipole term is O(N): set even when the pair sums are skipped, so
     the lag state of a replay starts where the reference's would */
  else if (p->use_ewald && p->dipole) t[ORC_T_DIPOLE] = dipole_energy_current(s);
  s->E_lj = t[ORC_T_LJ];
  s->E_real = t[ORC_T_REAL];
  s->E_kspace = t[ORC_T_KSPACE];
  s->E_self = t[ORC_T_SELF];
  s->E_dipole = t[ORC_T_DIPOLE];
  s->E_bond = t[ORC_T_BOND];
  s->E_wall = t[ORC_T_WALL];
  s->cur_dipl = s->trial_dipl = t[ORC_T_DIPOLE];
  for (int m = 0; m < n_mols; m++) s->bond_cur[m] = spring_molecule(s, m);
  return s;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  free(s->cur); free(s->trial); free(s->q); free(s->type); free(s->mol_start);
  free(s->mol_of); free(s->moved); free(s->bond_cur);
  free(s->kx); free(s->ky); free(s->kz); free(s->k2); free(s->ek2);
  free(s);
}

void orc_get_ewald_info(const orc_sys* s, orc_ewald_info* out) { *out = s->e; }
int orc_n_beads(const orc_sys* s) { return s->n; }
void orc_get_positions(const orc_sys* s, double* xyz) {
  memcpy(xyz, s->cur, sizeof(double) * 3 * s->n);
}

double orc_pair_lj(orc_sys* s, int i, int j) {
  return lj_pair_xyz(s, &s->trial[3 * i], s->type[i], &s->trial[3 * j], s->type[j]);
}
double orc_pair_real(orc_sys* s, int i, int j) {
  return real_pair_xyz(s, &s->trial[3 * i], s->q[i], &s->trial[3 * j], s->q[j]);
}
double orc_pair_repl(orc_sys* s, int i, int j) {
  return repl_pair_xyz(s, &s->trial[3 * i], s->q[i], &s->trial[3 * j], s->q[j]);
}
double orc_wall_bead(orc_sys* s, int i) {
  return wall_bead_z(s, s->trial[3 * i + 2], s->type[i]);
}

/* ------------------------------------------------------------------ */
/* full energy (ForceField::InitializeEnergy, force_field.cc:365-381)  */
/* ------------------------------------------------------------------ */
void orc_energy_full(orc_sys* s, double t[ORC_NTERMS]) {
  for (int k = 0; k < ORC_NTERMS; k++) t[k] = 0;
  const int* ms = s->mol_start;
  /* PotentialPair::EnergyInitialization (potential_pair.cc:55-100):
     intramolecular j<k, then intermolecular. */
  if (s->p.use_pair) {
    double e = 0;
    for (int m = 0; m < s->n_mols; m++)
      for (int j = ms[m]; j < ms[m + 1] - 1; j++)
        for (int k = j + 1; k < ms[m + 1]; k++) e += orc_pair_lj(s, j, k);
    for (int m1 = 0; m1 < s->n_mols - 1; m1++)
      for (int m2 = m1 + 1; m2 < s->n_mols; m2++)
        for (int k = ms[m1]; k < ms[m1 + 1]; k++)
          for (int l = ms[m2]; l < ms[m2 + 1]; l++) e += orc_pair_lj(s, k, l);
    t[ORC_T_LJ] = e;
  }
  /* PotentialEwald::EnergyInitialization (potential_ewald.cc:176-230) */
  if (s->p.use_ewald) {
    double er = 0, ek = 0, es = 0;
    for (int m1 = 0; m1 < s->n_mols; m1++)
      for (int m2 = m1; m2 < s->n_mols; m2++)
        for (int k = ms[m1]; k < ms[m1 + 1]; k++)
          for (int l = ms[m2]; l < ms[m2 + 1]; l++) {
            if ((m2 == m1 && l >= k) || m2 > m1) {
              double r = orc_pair_real(s, k, l);
              double g = orc_pair_repl(s, k, l);
              if (m2 == m1 && l == k) { r *= 0.5; g *= 0.5; }
              er += r;
              ek += g;
            }
          }
    for (int i = 0; i < s->n; i++) es += self_energy(s, s->q[i]);
    t[ORC_T_REAL] = er;
    t[ORC_T_KSPACE] = ek;
    t[ORC_T_SELF] = es;
    if (s->p.dipole) t[ORC_T_DIPOLE] = dipole_energy_current(s);
  }
  /* PotentialBond::EnergyInitialization (potential_bond.cc:16-27) */
  if (s->p.use_bond) {
    double e = 0;
    for (int m = 0; m < s->n_mols; m++) e += spring_molecule(s, m);
    t[ORC_T_BOND] = e;
  }
  /* PotentialExternal::EnergyInitialization (potential_external.cc:48-61) */
  if (s->p.use_ext) {
    double e = 0;
    for (int i = 0; i < s->n; i++) e += orc_wall_bead(s, i);
    t[ORC_T_WALL] = e;
  }
}

/* The same k-space total regrouped as (2 pi lB / V) sum_k ek2 |S(k)|^2
   over the full grid.  Used only to pin the pairwise form above. */
double orc_kspace_total_sk(orc_sys* s) {
  if (!s->p.use_ewald) return 0;
  double pref = s->p.lB / (kPi * s->e.vol_e) * (4 * kPi * kPi);
  double tot = 0;
  for (int ix = 0; ix < s->ceto[0]; ix++)
    for (int iy = 0; iy < s->ceto[1]; iy++)
      for (int iz = 0; iz < s->ceto[2]; iz++) {
        int idx = s->ceto[1] * s->ceto[2] * ix + s->ceto[2] * iy + iz;
        if (!(s->k2[idx] > 0 && s->k2[idx] <= s->e.repl_cutoff)) continue;
        double sr = 0, si = 0;
        for (int j = 0; j < s->n; j++) {
          const double* r = &s->trial[3 * j];
          double ph = s->kx[ix] * r[0] + s->ky[iy] * r[1] + s->kz[iz] * r[2];
          sr += s->q[j] * cos(ph);
          si += s->q[j] * sin(ph);
        }
        tot += s->ek2[idx] * (sr * sr + si * si);
      }
  return 0.5 * pref * tot;
}

/* ------------------------------------------------------------------ */
/* trial moves and energy differences                                  */
/* ------------------------------------------------------------------ */
void orc_set_trial(orc_sys* s, int mol, int n_moved, const int* idx,
                   const double* txyz) {
  (void)mol;
  for (int k = 0; k < n_moved; k++) {
    int i = idx[k];
    s->moved[i] = 1;
    for (int d = 0; d < 3; d++) s->trial[3 * i + d] = txyz[3 * k + d];
  }
}

int orc_get_moved(const orc_sys* s, int mol, int* idx, double* txyz) {
  int n = 0;
  for (int i = s->mol_start[mol]; i < s->mol_start[mol + 1]; i++)
    if (s->moved[i]) {
      if (idx) idx[n] = i;
      if (txyz) for (int d = 0; d < 3; d++) txyz[3 * n + d] = s->trial[3 * i + d];
      n++;
    }
  return n;
}

/* LJ pair on an explicit coordinate set: cur (c=0) or trial (c=1). */
static double lj_ij(orc_sys* s, int i, int ci, int j, int cj) {
  const double* a = ci ? &s->trial[3 * i] : &s->cur[3 * i];
  const double* b = cj ? &s->trial[3 * j] : &s->cur[3 * j];
  return lj_pair_xyz(s, a, s->type[i], b, s->type[j]);
}
static double real_ij(orc_sys* s, int i, int ci, int j, int cj) {
  const double* a = ci ? &s->trial[3 * i] : &s->cur[3 * i];
  const double* b = cj ? &s->trial[3 * j] : &s->cur[3 * j];
  return real_pair_xyz(s, a, s->q[i], b, s->q[j]);
}
static double repl_ij(orc_sys* s, int i, int ci, int j, int cj) {
  const double* a = ci ? &s->trial[3 * i] : &s->cur[3 * i];
  const double* b = cj ? &s->trial[3 * j] : &s->cur[3 * j];
  return repl_pair_xyz(s, a, s->q[i], b, s->q[j]);
}

/* PotentialPair::EnergyDifference (potential_pair.cc:152-201).  The cached
   old pair energy equals the pair function on current coordinates (it was
   computed from them and the beads have not moved since), so it is
   recomputed here rather than stored. */
static double pair_delta(orc_sys* s, int mol) {
  double dE = 0;
  const int a = s->mol_start[mol], b = s->mol_start[mol + 1];
  for (int i = a; i < b - 1; i++)
    for (int j = i + 1; j < b; j++)
      if (s->moved[i] || s->moved[j]) {
        double new_e = lj_ij(s, i, 1, j, 1);
        dE += (new_e - lj_ij(s, i, 0, j, 0));
      }
  for (int m = 0; m < s->n_mols; m++) {
    if (m == mol) continue;
    for (int j = s->mol_start[m]; j < s->mol_start[m + 1]; j++)
      for (int k = a; k < b; k++)
        if (s->moved[k]) {
          double new_e = lj_ij(s, k, 1, j, 1);
          dE += (new_e - lj_ij(s, k, 0, j, 0));
        }
  }
  return dE;
}

/* PotentialExternal::EnergyDifference (potential_external.cc:100-117) */
static double ext_delta(orc_sys* s, int mol) {
  double dE = 0;
  for (int i = s->mol_start[mol]; i < s->mol_start[mol + 1]; i++)
    if (s->moved[i]) {
      double eNew = wall_bead_z(s, s->trial[3 * i + 2], s->type[i]);
      if (eNew >= kVeryLargeEnergy) return kVeryLargeEnergy;
      dE += eNew - wall_bead_z(s, s->cur[3 * i + 2], s->type[i]);
    }
  return dE;
}

/* PotentialEwald::EnergyDifference (potential_ewald.cc:429-535).  Returns the
   total in the reference's summation order and the per-term split. */
static double ewald_delta(orc_sys* s, int mol, double* d_real, double* d_repl,
                          double* d_self, double* d_dip) {
  double dE = 0, dr = 0, dk = 0, ds = 0;
  const int a = s->mol_start[mol], b = s->mol_start[mol + 1];
  for (int i = a; i < b; i++)
    for (int j = i; j < b; j++)
      if (s->moved[i] || s->moved[j]) {
        double nr = real_ij(s, i, 1, j, 1), nk = repl_ij(s, i, 1, j, 1);
        double orl = real_ij(s, i, 0, j, 0), ok = repl_ij(s, i, 0, j, 0);
        if (j == i) { nr *= 0.5; nk *= 0.5; orl *= 0.5; ok *= 0.5; }
        dE += (nr + nk - (orl + ok));
        dr += nr - orl;
        dk += nk - ok;
      }
  /* moved beads x the other molecules: the pair values are evaluated in
     parallel into a buffer (OpenMP over the partner beads) and then summed
     serially in the reference's loop order, so the result is the same,
     bit for bit, as the serial loop's */
  int nmv = 0;
  for (int k = a; k < b; k++) nmv += s->moved[k] ? 1 : 0;
  const int n_out = s->n - (b - a);
  if (nmv > 0 && n_out > 0) {
    int* mk = (int*)malloc(sizeof(int) * nmv);
    double* pv = (double*)malloc(sizeof(double) * 4 * (size_t)nmv * n_out);
    int c = 0;
    for (int k = a; k < b; k++)
      if (s->moved[k]) mk[c++] = k;
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < n_out; jj++) {
      const int j = jj < a ? jj : jj + (b - a);
      for (int u = 0; u < nmv; u++) {
        double* o = &pv[4 * ((size_t)jj * nmv + u)];
        o[0] = real_ij(s, mk[u], 1, j, 1);
        o[1] = repl_ij(s, mk[u], 1, j, 1);
        o[2] = real_ij(s, mk[u], 0, j, 0);
        o[3] = repl_ij(s, mk[u], 0, j, 0);
      }
    }
    /* molecule order m = 0..n_mols-1 without mol is bead order without
       [a, b), since molecules are contiguous and ordered */
    for (int jj = 0; jj < n_out; jj++)
      for (int u = 0; u < nmv; u++) {
        const double* o = &pv[4 * ((size_t)jj * nmv + u)];
        double nr = o[0], nk = o[1], orl = o[2], ok = o[3];
        dE += (nr + nk - (orl + ok));
        dr += nr - orl;
        dk += nk - ok;
      }
    free(pv);
    free(mk);
  }
  for (int i = a; i < b; i++)
    if (s->moved[i]) {
      double e = self_energy(s, s->q[i]) - self_energy(s, s->q[i]);
      dE += e;
      ds += e;
    }
  double dd = 0;
  if (s->p.dipole) {
    /* :528-531 -- evaluated on CURRENT coordinates (the lag) */
    s->trial_dipl = dipole_energy_current(s);
    dd = s->trial_dipl - s->cur_dipl;
    dE += dd;
  }
  *d_real = dr; *d_repl = dk; *d_self = ds; *d_dip = dd;
  return dE;
}

/* ForceField::EnergyDifference (force_field.cc:407-434) */
double orc_energy_difference(orc_sys* s, int mol, double t[ORC_NTERMS]) {
  for (int k = 0; k < ORC_NTERMS; k++) t[k] = 0;
  double dE = 0;
  s->pend_valid = 1;
  if (s->p.use_pair) {
    t[ORC_T_LJ] = pair_delta(s, mol);
    dE += t[ORC_T_LJ];
    if (dE >= kVeryLargeEnergy) goto done;
  }
  if (s->p.use_ext) {
    t[ORC_T_WALL] = ext_delta(s, mol);
    dE += t[ORC_T_WALL];
    if (dE >= kVeryLargeEnergy) goto done;
  }
  if (s->p.use_ewald) {
    double dr, dk, ds, dd;
    dE += ewald_delta(s, mol, &dr, &dk, &ds, &dd);
    t[ORC_T_REAL] = dr; t[ORC_T_KSPACE] = dk; t[ORC_T_SELF] = ds; t[ORC_T_DIPOLE] = dd;
  }
  if (s->p.use_bond) {
    /* PotentialSpring::EnergyDifference (potential_spring.cc:63-80) */
    if (s->mol_start[mol + 1] - s->mol_start[mol] > 1) {
      double te = spring_molecule(s, mol);
      t[ORC_T_BOND] = te - s->bond_cur[mol];
      dE += t[ORC_T_BOND];
    }
  }
done:
  for (int k = 0; k < ORC_NTERMS; k++) s->pend[k] = t[k];
  return dE;
}

/* ForceField::FinalizeEnergies (force_field.cc:436-451) followed by the
   commit/rollback and UnsetMoved of Simulation::TranslationalMove
   (simulation.cc:334-351). */
void orc_finalize(orc_sys* s, int mol, int accept) {
  const int a = s->mol_start[mol], b = s->mol_start[mol + 1];
  if (accept && s->pend_valid) {
    s->E_lj += s->pend[ORC_T_LJ];
    s->E_real += s->pend[ORC_T_REAL];
    s->E_kspace += s->pend[ORC_T_KSPACE];
    s->E_self += s->pend[ORC_T_SELF];
    s->E_dipole += s->pend[ORC_T_DIPOLE];
    s->E_bond += s->pend[ORC_T_BOND];
    s->E_wall += s->pend[ORC_T_WALL];
  }
  /* potential_ewald.cc:614-622 */
  if (s->p.dipole) {
    if (accept) s->cur_dipl = s->trial_dipl;
    else s->trial_dipl = s->cur_dipl;
  }
  for (int i = a; i < b; i++) {
    for (int d = 0; d < 3; d++) {
      if (accept) s->cur[3 * i + d] = s->trial[3 * i + d];
      else s->trial[3 * i + d] = s->cur[3 * i + d];
    }
    s->moved[i] = 0;
  }
  /* potential_bond.cc:69-82: the trial molecule energy becomes current */
  if (accept && s->p.use_bond && b - a > 1) s->bond_cur[mol] = spring_molecule(s, mol);
  s->pend_valid = 0;
}

void orc_running_totals(const orc_sys* s, double t[ORC_NTERMS]) {
  t[ORC_T_LJ] = s->E_lj; t[ORC_T_REAL] = s->E_real; t[ORC_T_KSPACE] = s->E_kspace;
  t[ORC_T_SELF] = s->E_self; t[ORC_T_DIPOLE] = s->E_dipole; t[ORC_T_BOND] = s->E_bond;
  t[ORC_T_WALL] = s->E_wall;
}

/* ------------------------------------------------------------------ */
/* std::mt19937 (C++ [rand.predef]; Matsumoto & Nishimura 1998)         */
/* ------------------------------------------------------------------ */
void orc_mt_seed(orc_mt19937* r, uint32_t seed) {
  r->mt[0] = seed;
  for (int i = 1; i < 624; i++)
    r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
  r->mti = 624;
}

uint32_t orc_mt_next(orc_mt19937* r) {
  if (r->mti >= 624) {
    for (int i = 0; i < 624; i++) {
      uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = r->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      r->mt[i] = v;
    }
    r->mti = 0;
  }
  uint32_t y = r->mt[r->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* (double)rand_gen()/rand_gen.max() */
static double u01(orc_mt19937* r) { return (double)orc_mt_next(r) / 4294967295.0; }

/* randSphere (misc.cc:95-109), Marsaglia 1972 */
void orc_rand_sphere(orc_mt19937* r, double vec[3]) {
  double rand_square = 2, r1 = 0, r2 = 0;
  while (rand_square > 1) {
    r1 = 1 - 2 * u01(r);
    r2 = 1 - 2 * u01(r);
    rand_square = r1 * r1 + r2 * r2;
  }
  double ranh = 2 * sqrt(1 - rand_square);
  vec[0] = r1 * ranh;
  vec[1] = r2 * ranh;
  vec[2] = (1 - 2 * rand_square);
}

/* ------------------------------------------------------------------ */
/* move generators (molecule.cc:103-312)                               */
/* ------------------------------------------------------------------ */
static int count_moved(const orc_sys* s, int mol) {
  int n = 0;
  for (int i = s->mol_start[mol]; i < s->mol_start[mol + 1]; i++) n += s->moved[i];
  return n;
}

/* Molecule::BeadTranslate (molecule.cc:103-133) */
int orc_move_bead_translate(orc_sys* s, int mol, double move_size, orc_mt19937* r) {
  double vec[3];
  orc_rand_sphere(r, vec);
  double vec_len = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
  if (vec_len > 0) vec_len = 3 * move_size / vec_len;
  int b0 = s->mol_start[mol];
  s->moved[b0] = 1;
  for (int i = 0; i < 3; i++) s->trial[3 * b0 + i] = s->cur[3 * b0 + i] + vec_len * vec[i];
  return count_moved(s, mol);
}

/* Molecule::COMTranslate (molecule.cc:135-151) */
int orc_move_com(orc_sys* s, int mol, double move_size, orc_mt19937* r) {
  double v0 = 0.5 * move_size * u01(r);
  double v1 = 0.5 * move_size * u01(r);
  double v2 = 0.5 * move_size * u01(r);
  double vec[3] = {v0, v1, v2};
  if (orc_mt_next(r) % 2 == 0) vec[0] = -vec[0];
  if (orc_mt_next(r) % 2 == 0) vec[1] = -vec[1];
  if (orc_mt_next(r) % 2 == 0) vec[2] = -vec[2];
  for (int i = s->mol_start[mol]; i < s->mol_start[mol + 1]; i++) {
    for (int j = 0; j < 3; j++) s->trial[3 * i + j] = s->cur[3 * i + j] + vec[j];
    s->moved[i] = 1;
  }
  return count_moved(s, mol);
}

/* Molecule::Pivot (molecule.cc:153-237) */
int orc_move_pivot(orc_sys* s, int mol, double move_size, orc_mt19937* r,
                   double eq_bond_len, int vary_bond) {
  const int base = s->mol_start[mol];
  const int len = s->mol_start[mol + 1] - base;
  double* T = &s->trial[3 * base];
  int pivot = (int)floor(len * u01(r));
  if (pivot == len) pivot--;
  double msr = move_size * u01(r);
  for (int i = pivot + 1; i < len; i++) {
    double vec[3];
    orc_rand_sphere(r, vec);
    double x = T[3 * i + 0] + msr * vec[0];
    double y = T[3 * i + 1] + msr * vec[1];
    double z = T[3 * i + 2] + msr * vec[2];
    double dx = x - T[3 * (i - 1) + 0], dy = y - T[3 * (i - 1) + 1], dz = z - T[3 * (i - 1) + 2];
    double bond_len = eq_bond_len;
    if (vary_bond) bond_len += (eq_bond_len / 5.0) * (u01(r) - 0.5);
    double norm = bond_len / sqrt(dx * dx + dy * dy + dz * dz);
    double cx = norm * dx + T[3 * (i - 1) + 0];
    double cy = norm * dy + T[3 * (i - 1) + 1];
    double cz = norm * dz + T[3 * (i - 1) + 2];
    double mx = cx - T[3 * i + 0], my = cy - T[3 * i + 1], mz = cz - T[3 * i + 2];
    for (int j = i; j < len; j++) {
      T[3 * j + 0] = T[3 * j + 0] + mx;
      T[3 * j + 1] = T[3 * j + 1] + my;
      T[3 * j + 2] = T[3 * j + 2] + mz;
    }
  }
  for (int i = pivot - 1; i >= 0; i--) {
    double vec[3];
    orc_rand_sphere(r, vec);
    double x = T[3 * i + 0] + msr * vec[0];
    double y = T[3 * i + 1] + msr * vec[1];
    double z = T[3 * i + 2] + msr * vec[2];
    double dx = x - T[3 * (i + 1) + 0], dy = y - T[3 * (i + 1) + 1], dz = z - T[3 * (i + 1) + 2];
    double bond_len = eq_bond_len;
    if (vary_bond) bond_len += (eq_bond_len / 5.0) * (u01(r) - 0.5);
    double norm = bond_len / sqrt(dx * dx + dy * dy + dz * dz);
    double cx = norm * dx + T[3 * (i + 1) + 0];
    double cy = norm * dy + T[3 * (i + 1) + 1];
    double cz = norm * dz + T[3 * (i + 1) + 2];
    double mx = cx - T[3 * i + 0], my = cy - T[3 * i + 1], mz = cz - T[3 * i + 2];
    for (int j = i; j >= 0; j--) {
      T[3 * j + 0] = T[3 * j + 0] + mx;
      T[3 * j + 1] = T[3 * j + 1] + my;
      T[3 * j + 2] = T[3 * j + 2] + mz;
    }
  }
  for (int i = 0; i < len; i++) s->moved[base + i] = 1;
  return count_moved(s, mol);
}

/* Molecule::Crankshaft (molecule.cc:239-265).  Eigen's AngleAxisd rotation
   is restated as the Rodrigues matrix R = c I + s [a]x + (1-c) a a^T. */
int orc_move_crankshaft(orc_sys* s, int mol, double move_size, orc_mt19937* r) {
  const int base = s->mol_start[mol];
  const int len = s->mol_start[mol + 1] - base;
  int first = (int)floor((len - 2) * u01(r));
  int last = (int)floor((len - first) * u01(r)) + first;
  if (last >= len) last = len - 1;  /* out-of-range guard (p = 2^-32) */
  const double* C = &s->cur[3 * base];
  double* T = &s->trial[3 * base];
  double ax[3];
  for (int i = 0; i < 3; i++) ax[i] = C[3 * last + i] - C[3 * first + i];
  double nn = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
  if (nn > 0) {
    double inv = sqrt(nn);
    for (int i = 0; i < 3; i++) ax[i] /= inv;
  }
  double angle = move_size * (u01(r) * 2 * M_PI - M_PI);
  double sn = sin(angle), c = cos(angle);
  double R[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = (1 - c) * ax[i] * ax[j] + (i == j ? c : 0);
  R[0][1] -= sn * ax[2]; R[1][0] += sn * ax[2];
  R[0][2] += sn * ax[1]; R[2][0] -= sn * ax[1];
  R[1][2] -= sn * ax[0]; R[2][1] += sn * ax[0];
  for (int i = first + 1; i < last; i++) {
    s->moved[base + i] = 1;
    double v[3];
    for (int k = 0; k < 3; k++) v[k] = C[3 * i + k] - C[3 * first + k];
    for (int k = 0; k < 3; k++)
      T[3 * i + k] = R[k][0] * v[0] + R[k][1] * v[1] + R[k][2] * v[2] + C[3 * first + k];
  }
  return count_moved(s, mol);
}

/* Molecule::RandomReptation (molecule.cc:267-312) */
int orc_move_reptation(orc_sys* s, int mol, orc_mt19937* r, double eq_bond_len,
                       int vary_bond) {
  const int base = s->mol_start[mol];
  const int len = s->mol_start[mol + 1] - base;
  if (eq_bond_len < 0) return -1;
  double bond_len = eq_bond_len;
  if (vary_bond) bond_len += (eq_bond_len / 5.0) * (u01(r) - 0.5);
  int direction = 1, begin = 0, end = len - 1;
  if (orc_mt_next(r) % 2 == 0) { direction = -1; begin = len - 1; end = 0; }
  const double* C = &s->cur[3 * base];
  double* T = &s->trial[3 * base];
  for (int i = begin; i != end; i += direction)
    for (int j = 0; j < 3; j++) T[3 * i + j] = C[3 * (i + direction) + j];
  double vec[3];
  orc_rand_sphere(r, vec);
  double vec_len = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
  for (int i = 0; i < 3; i++) T[3 * end + i] = C[3 * end + i] + bond_len * vec[i] / vec_len;
  for (int i = 0; i < len; i++) s->moved[base + i] = 1;
  return count_moved(s, mol);
}

/* ------------------------------------------------------------------ */
/* Simulation::Run step + TranslationalMove (simulation.cc:216-355)     */
/* ------------------------------------------------------------------ */
void orc_mc_step(orc_sys* s, orc_mt19937* r, const double move_prob[5],
                 double move_size, int phantom, int grafted, double bond_len,
                 int vary_bond, orc_step_info* info) {
  (void)orc_mt_next(r);  /* simulation.cc:221, the per-step draw */
  orc_translational_move(s, r, move_prob, move_size, phantom, grafted, bond_len, vary_bond, info);
}

void orc_translational_move(orc_sys* s, orc_mt19937* r, const double move_prob[5],
                            double move_size, int phantom, int grafted, double bond_len,
                            int vary_bond, orc_step_info* info) {
  memset(info, 0, sizeof(*info));
  info->mol = -1;
  int n_mol = s->n_mols;
  /* Simulation::UpdateMolCounts (simulation.cc:1066-1094) */
  int n_chain = 0, n_cion = 0, n_aion = 0, n_nion = 0;
  for (int i = phantom; i < n_mol; i++) {
    int sz = s->mol_start[i + 1] - s->mol_start[i];
    double q0 = s->q[s->mol_start[i]];
    if (sz > 1) n_chain++;
    else if (q0 > 0) n_cion++;
    else if (q0 < 0) n_aion++;
    else n_nion++;
  }
  n_chain -= grafted;
  if (n_mol - phantom <= 0) return;
  int chain_id = -1, ion_id = -1;
  int n_ions = n_cion + n_aion + n_nion;
  int which_chain = (int)floor(u01(r) * (grafted + n_chain));
  int which_ion = (int)floor(u01(r) * n_ions);
  if (which_chain == grafted + n_chain) which_chain--;
  if (which_ion == n_ions) which_ion--;
  int cc = -1, ic = -1;
  for (int i = phantom; i < n_mol; i++) {
    int sz = s->mol_start[i + 1] - s->mol_start[i];
    if (sz > 1) cc++; else ic++;
    if (cc == which_chain) { chain_id = i; which_chain = -2; }
    if (ic == which_ion) { ion_id = i; which_ion = -2; }
  }
  int move_type = 0;
  double rn = u01(r);
  double cur_type = move_prob[0];
  while (cur_type < rn) { move_type++; cur_type += move_prob[move_type]; }
  if (move_type >= 5) move_type = 4;
  int mol_id = -1;
  if (move_type == 0 && n_ions > 0) {
    mol_id = ion_id;
    orc_move_bead_translate(s, mol_id, move_size, r);
  } else if (grafted + n_chain > 0) {
    mol_id = chain_id;
    switch (move_type) {
      case 1: orc_move_com(s, mol_id, move_size, r); break;
      case 2: orc_move_pivot(s, mol_id, move_size, r, bond_len, vary_bond); break;
      case 3: orc_move_crankshaft(s, mol_id, move_size, r); break;
      case 4: orc_move_reptation(s, mol_id, r, bond_len, vary_bond); break;
    }
  }
  if ((move_type == 0 && n_ions > 0) || (move_type > 0 && grafted + n_chain > 0)) {
    info->attempted = 1;
    info->move_type = move_type;
    info->mol = mol_id;
    info->n_moved = count_moved(s, mol_id);
    double dE = orc_energy_difference(s, mol_id, info->terms);
    int accept = 0;
    if (dE >= kVeryLargeEnergy) accept = 0;
    else accept = (u01(r) < exp(-s->p.beta * dE));
    info->dE = dE;
    info->accepted = accept;
    orc_finalize(s, mol_id, accept);
  }
}

double orc_bead_energy_sample(orc_sys* s, const double xyz[3], double q, int t,
                              int n_partners) {
  double e = 0;
  if (n_partners > s->n) n_partners = s->n;
  for (int j = 0; j < n_partners; j++) {
    const double* b = &s->trial[3 * j];
    if (s->p.use_pair) e += lj_pair_xyz(s, xyz, t, b, s->type[j]);
    if (s->p.use_ewald) {
      e += real_pair_xyz(s, xyz, q, b, s->q[j]);
      e += repl_pair_xyz(s, xyz, q, b, s->q[j]);
    }
  }
  return e;
}

/* PotentialEwaldCoul::DipoleEDiff (potential_ewald_coul.cc:473-530, called at
   cbmc.cc:109-112): with M_z the sum of q z over the molecules outside
   [ex_lo, ex_hi] (CURRENT z), then the grown chain beads, then their
   counterions, the change 2 pi l_B / V (M_n^2 - M_z^2) when the trial bead
   (q1 at b1) and, for a charged chain, its counterion (q2 at b2) join:
   M_n = M_z + q1 z1 (+ q2 z2). */
double orc_cbmc_dipole_ediff(const orc_sys* s, const double b1[3], double q1, const double b2[3], double q2,
                             int charged, const double* grown_c, const double* grown_i, int cur, int ex_lo,
                             int ex_hi) {
  double mo = 0;
  for (int m = 0; m < s->n_mols; m++) {
    if (ex_lo <= ex_hi && m >= ex_lo && m <= ex_hi) continue;
    for (int j = s->mol_start[m]; j < s->mol_start[m + 1]; j++) mo += s->q[j] * s->cur[3 * j + 2];
  }
  for (int i = 0; i < cur; i++) mo += q1 * grown_c[3 * i + 2];
  if (charged)
    for (int i = 0; i < cur; i++) mo += q2 * grown_i[3 * i + 2];
  double mn = mo + q1 * b1[2];
  if (charged) mn += q2 * b2[2];
  return s->p.lB * 2 * kPi / (s->e.vol_e / 1.0) * (mn * mn - mo * mo);
}

/* ForceField::BeadsEnergy (cbmc.cc:5-151): trial chain bead b1 (type t1,
   charge q1) and, for charged chains, its trial counterion b2, against the
   molecules outside [ex_lo, ex_hi] (cbmc.cc:23; none when ex_lo > ex_hi),
   the grown chain beads grown_c[0..cur) and counterions grown_i[0..cur),
   with the reference's break on LJ >= 1e8, and the slab dipole term
   (orc_cbmc_dipole_ediff) with the self energies.  Trial coordinates as the
   reference reads them (equal to the current ones outside a pending move). */
double orc_cbmc_beads_energy(orc_sys* s, const double b1[3], int t1, double q1, const double b2[3],
                             int t2, double q2, int charged, const double* grown_c,
                             const double* grown_i, int cur, int ex_lo, int ex_hi) {
  double pair_e = 0, ewald_e = 0;
  double pair_e1 = 0, pair_e2 = 0, ewald_r1 = 0, ewald_r2 = 0, ewald_k1 = 0, ewald_k2 = 0;
  const int use_pair = s->p.use_pair, use_ewald = s->p.use_ewald;
  /* 1. other molecules */
  for (int m = 0; m < s->n_mols; m++) {
    if (ex_lo <= ex_hi && m >= ex_lo && m <= ex_hi) continue;
    for (int j = s->mol_start[m]; j < s->mol_start[m + 1]; j++) {
      const double* bj = &s->trial[3 * j];
      if (use_pair) {
        pair_e1 = lj_pair_xyz(s, b1, t1, bj, s->type[j]);
        if (charged) pair_e2 = lj_pair_xyz(s, b2, t2, bj, s->type[j]);
      }
      pair_e += pair_e1 + pair_e2;
      if (pair_e >= kVeryLargeEnergy) break;
      if (use_ewald) {
        ewald_r1 = real_pair_xyz(s, b1, q1, bj, s->q[j]);
        ewald_k1 = repl_pair_xyz(s, b1, q1, bj, s->q[j]);
        if (charged) {
          ewald_r2 = real_pair_xyz(s, b2, q2, bj, s->q[j]);
          ewald_k2 = repl_pair_xyz(s, b2, q2, bj, s->q[j]);
        }
      }
      ewald_e += ewald_r1 + ewald_k1 + ewald_r2 + ewald_k2;
    }
    if (pair_e >= kVeryLargeEnergy) break;
  }
  /* 2. the grown part of the trial chain */
  for (int i = 0; i < cur; i++) {
    if (use_pair) {
      pair_e1 = 0;
      if (i < cur - 1) pair_e1 += lj_pair_xyz(s, b1, t1, &grown_c[3 * i], t1);
      if (charged) {
        pair_e1 += lj_pair_xyz(s, b1, t1, &grown_i[3 * i], t2);
        pair_e2 = lj_pair_xyz(s, b2, t2, &grown_c[3 * i], t1);
        pair_e2 += lj_pair_xyz(s, b2, t2, &grown_i[3 * i], t2);
      }
    }
    pair_e += pair_e1 + pair_e2;
    if (pair_e >= kVeryLargeEnergy) break;
    if (use_ewald) {
      ewald_r1 = real_pair_xyz(s, b1, q1, &grown_c[3 * i], q1);
      ewald_k1 = repl_pair_xyz(s, b1, q1, &grown_c[3 * i], q1);
      if (charged) {
        ewald_r1 += real_pair_xyz(s, b1, q1, &grown_i[3 * i], q2);
        ewald_k1 += repl_pair_xyz(s, b1, q1, &grown_i[3 * i], q2);
        ewald_r2 = real_pair_xyz(s, b2, q2, &grown_c[3 * i], q1);
        ewald_k2 = repl_pair_xyz(s, b2, q2, &grown_c[3 * i], q1);
        ewald_r2 += real_pair_xyz(s, b2, q2, &grown_i[3 * i], q2);
        ewald_k2 += repl_pair_xyz(s, b2, q2, &grown_i[3 * i], q2);
      }
    }
    ewald_e += ewald_r1 + ewald_k1 + ewald_r2 + ewald_k2;
  }
  /* 3. the pair (b1, b2), self energies and self images */
  if (use_pair && charged) pair_e += lj_pair_xyz(s, b1, t1, b2, t2);
  if (use_ewald && pair_e < kVeryLargeEnergy) {
    ewald_e += self_energy(s, q1);
    ewald_e += 0.5 * real_pair_xyz(s, b1, q1, b1, q1);
    ewald_e += 0.5 * repl_pair_xyz(s, b1, q1, b1, q1);
    if (charged) {
      ewald_e += self_energy(s, q2);
      ewald_e += real_pair_xyz(s, b1, q1, b2, q2);
      ewald_e += repl_pair_xyz(s, b1, q1, b2, q2);
      ewald_e += 0.5 * real_pair_xyz(s, b2, q2, b2, q2);
      ewald_e += 0.5 * repl_pair_xyz(s, b2, q2, b2, q2);
    }
    if (s->p.dipole)
      ewald_e += orc_cbmc_dipole_ediff(s, b1, q1, b2, q2, charged, grown_c, grown_i, cur, ex_lo, ex_hi);
  }
  /* 4. walls */
  if (s->p.use_ext) {
    pair_e += wall_bead_z(s, b1[2], t1);
    if (charged) pair_e += wall_bead_z(s, b2[2], t2);
  }
  if (pair_e >= kVeryLargeEnergy) return kVeryLargeEnergy;
  return pair_e + ewald_e;
}

/* Simulation::RadiusOfGyration (simulation.cc:406-434) over all chains. */
double orc_radius_of_gyration(const orc_sys* s) {
  static const double nobox[3] = {0, 0, 0};
  double rg = 0;
  int nc = 0;
  for (int m = 0; m < s->n_mols; m++) {
    int a = s->mol_start[m], b = s->mol_start[m + 1], len = b - a;
    if (len <= 1) continue;
    double rm = 0;
    for (int j = a; j < b - 1; j++)
      for (int k = j + 1; k < b; k++) {
        double r = bb_dist(&s->cur[3 * j], &s->cur[3 * k], nobox, 0);
        rm += r * r;
      }
    rm /= (len * len);
    rg += rm;
    nc++;
  }
  return nc ? rg / nc : 0;
}

/* Simulation::EndToEndDistance (simulation.cc:483-501): mean squared. */
double orc_end_to_end(const orc_sys* s) {
  static const double nobox[3] = {0, 0, 0};
  double e = 0;
  int nc = 0;
  for (int m = 0; m < s->n_mols; m++) {
    int a = s->mol_start[m], b = s->mol_start[m + 1];
    if (b - a <= 1) continue;
    double d = bb_dist(&s->cur[3 * a], &s->cur[3 * (b - 1)], nobox, 0);
    e += d * d;
    nc++;
  }
  return nc ? e / nc : 0;
}

/* ------------------------------------------------------------------ */
/* grand-canonical chain exchange (cbmc.cc:152-533)                    */
/* ------------------------------------------------------------------ */

/* gasdev (misc.cc:111-133): Marsaglia's polar method, the second deviate
   of each pair cached for the next call. */
double orc_gasdev(double mean, double stdev, orc_mt19937* r, orc_gauss* gs) {
  if (gs->iset == 0) {
    double v1, v2, rsq;
    do {
      v1 = 2.0 * u01(r) - 1.0;
      v2 = 2.0 * u01(r) - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    const double fac = sqrt(-2.0 * log(rsq) / rsq);
    gs->gset = v1 * fac;
    gs->iset = 1;
    return v2 * fac * stdev + mean;
  }
  gs->iset = 0;
  return gs->gset * stdev + mean;
}

/* PotentialSpring::RandomBondLen (potential_spring.cc:50-61): Gaussian
   around r0 of width 1/sqrt(beta k), thinned by len^2 / (r0 + 3 sigma)^2. */
static double random_bond_len(const orc_sys* s, orc_mt19937* r, orc_gauss* gs) {
  const double sigma = sqrt(1 / (s->p.beta * s->p.k_bond));
  const double a = (s->p.r0 + 3 * sigma) * (s->p.r0 + 3 * sigma);
  double len = 0;
  int ready = 0;
  while (!ready) {
    len = orc_gasdev(s->p.r0, sigma, r, gs);
    ready = u01(r) < (len * len / a);
  }
  return len;
}

double orc_random_bond_len(orc_sys* s, orc_mt19937* r, orc_gauss* gs) {
  return random_bond_len(s, r, gs);
}

/* ForceField::UpdateMolCounts (force_field.cc:1258-1277) */
static void gc_counts(const orc_sys* s, const orc_gc* gc, int* n_chain, int* n_cion, int* n_aion) {
  int c = 0, ci = 0, ai = 0;
  for (int m = gc->phantom; m < s->n_mols; m++) {
    const int a = s->mol_start[m], len = s->mol_start[m + 1] - a;
    if (len > 1) c++;
    else if (s->q[a] >= 0) ci++;
    else ai++;
  }
  *n_chain = c - gc->grafted;
  *n_cion = ci;
  *n_aion = ai;
}

static int gc_spc1(const orc_gc* gc, int n_chain, int n_cion, int n_aion) {
  if (gc->chain_len > 1) return n_chain;
  if (gc->bead_charge >= 0) return n_cion - gc->coion;
  return n_aion - gc->coion;
}

static int gc_spc2(const orc_gc* gc, int n_cion, int n_aion) {
  if (gc->bead_charge == 0) return 0;
  return gc->bead_charge > 0 ? n_aion : n_cion;
}

/* ForceField::CBMCFGenTrialBeads (cbmc.cc:161-212): K trial chain beads at a
   sampled bond length in a random direction from `end`, each with a
   counterion uniform in the box for a charged chain; tri holds the chain
   trials then the counterion trials.  Returns W = sum of the weights. */
static double gen_trials(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, const double* end,
                         const double* chain, int cur, int ex_lo, int ex_hi, double* tri, double* wts) {
  const int K = gc->trials, L = gc->chain_len, ch = gc->bead_charge != 0;
  double W = 0;
  for (int i = 0; i < K; i++) {
    double bond_len = 0;
    if (gc->bond == 1) bond_len = random_bond_len(s, r, gs);
    else if (gc->bond == 2) bond_len = gc->rigid_bond;
    double v[3];
    orc_rand_sphere(r, v);
    for (int j = 0; j < 3; j++) tri[3 * i + j] = v[j] * bond_len + end[j];
    if (ch)
      for (int j = 0; j < 3; j++) tri[3 * (K + i) + j] = u01(r) * s->p.box[j];
    const double* b2 = ch ? &tri[3 * (K + i)] : &tri[3 * i];
    const double e = orc_cbmc_beads_energy(s, &tri[3 * i], gc->type_c, gc->bead_charge, b2, gc->type_c,
                                           -gc->bead_charge, ch, chain, chain + 3 * L, cur, ex_lo, ex_hi);
    wts[i] = exp(-s->p.beta * e);
    W += wts[i];
  }
  return W;
}

/* The growth of a new chain shared by the insertion and the test insertion
   of the chemical potential: first bead(s) uniform in the box, the rest
   chosen among the trials by their weights.  Returns the Rosenbluth weight;
   widom selects the chemical-potential variant (no early return after the
   first bead, cbmc.cc:489-500). */
static double grow_chain(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double* chain, int widom,
                         int first_given) {
  const int K = gc->trials, L = gc->chain_len, ch = gc->bead_charge != 0;
  double* tri = (double*)malloc(sizeof(double) * 6 * K);
  double* wts = (double*)malloc(sizeof(double) * K);
  double weight = 1.0;
  if (!first_given)
    for (int i = 0; i < (ch ? 2 : 1); i++)
      for (int j = 0; j < 3; j++) chain[3 * (i * L) + j] = u01(r) * s->p.box[j];
  const double e0 = orc_cbmc_beads_energy(s, chain, gc->type_c, gc->bead_charge, ch ? &chain[3 * L] : chain,
                                          gc->type_c, -gc->bead_charge, ch, chain, chain + 3 * L, 0, 0, -1);
  if (widom) {
    if (e0 == 1e8) weight *= 0;
    else weight *= exp(-s->p.beta * e0);
  } else {
    weight *= exp(-s->p.beta * e0);
    if (weight <= 0) goto done;
  }
  for (int i = 1; i < L; i++) {
    const double Wi = gen_trials(s, r, gs, gc, &chain[3 * (i - 1)], chain, i, 0, -1, tri, wts);
    weight *= Wi / K;
    if (weight <= 0) break;
    const double rn = u01(r) * Wi;
    int cb = 0;
    double cum = wts[0];
    while (cum < rn && cb < K - 1) {
      cb++;
      cum += wts[cb];
    }
    for (int j = 0; j < 3; j++) chain[3 * i + j] = tri[3 * cb + j];
    if (ch)
      for (int j = 0; j < 3; j++) chain[3 * (i + L) + j] = tri[3 * (K + cb) + j];
  }
done:
  free(tri);
  free(wts);
  return weight;
}

/* The insertion's prefactor C of cbmc.cc:268-300. */
static double insertion_factor(orc_sys* s, const orc_gc* gc) {
  const int L = gc->chain_len;
  int n_chain, n_cion, n_aion;
  gc_counts(s, gc, &n_chain, &n_cion, &n_aion);
  const int spc1 = gc_spc1(gc, n_chain, n_cion, n_aion), spc2 = gc_spc2(gc, n_cion, n_aion);
  const int spc1_add = 1, spc2_add = gc->bead_charge == 0 ? 0 : L;
  double factorial = 1;
  for (int i = spc1 + 1; i <= spc1 + spc1_add; i++) factorial *= i;
  for (int i = spc2 + 1; i <= spc2 + spc2_add; i++) factorial *= i;
  const double vol = s->p.box[0] * s->p.box[1] * s->p.box[2];
  double vol_over_lam = 1;
  vol_over_lam *= pow(vol / (gc->debroglie / pow((double)L, 1.5)), spc1_add);
  vol_over_lam *= pow(vol / (gc->debroglie / pow(1.0, 1.5)), spc2_add);
  return vol_over_lam * (1.0 / factorial);
}

/* The deletion's prefactor C of cbmc.cc:403-441 (counts of the current
   system, before the deletion). */
static double deletion_factor(orc_sys* s, const orc_gc* gc) {
  const int L = gc->chain_len;
  int n_chain, n_cion, n_aion;
  gc_counts(s, gc, &n_chain, &n_cion, &n_aion);
  const int spc1 = gc_spc1(gc, n_chain, n_cion, n_aion), spc2 = gc_spc2(gc, n_cion, n_aion);
  const int spc1_del = 1, spc2_del = gc->bead_charge == 0 ? 0 : L;
  double factorial = 1;
  for (int i = spc1; i > spc1 - spc1_del; i--) factorial *= i;
  for (int i = spc2; i > spc2 - spc2_del; i--) factorial *= i;
  const double vol = s->p.box[0] * s->p.box[1] * s->p.box[2];
  double lam_over_vol = 1;
  lam_over_vol *= pow((gc->debroglie / pow((double)L, 1.5)) / vol, spc1_del);
  lam_over_vol *= pow((gc->debroglie / pow(1.0, 1.5)) / vol, spc2_del);
  return lam_over_vol * factorial;
}

int orc_gc_insertion(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double* chain_out,
                     double* weight_out) {
  const int L = gc->chain_len;
  memset(chain_out, 0, sizeof(double) * 6 * L);
  const double weight = grow_chain(s, r, gs, gc, chain_out, 0, 0);
  if (weight_out) *weight_out = weight;
  if (weight <= 0) return 0;
  const double C = insertion_factor(s, gc);
  const double rn = u01(r);
  return rn < (exp(s->p.beta * gc->chem_pot) * weight) * C;
}

/* Boltzmann factor of a first chain bead b1 and counterion b2 against the
   molecules outside [ex_lo, ex_hi] (BeadsEnergy with nothing grown). */
static double first_bead_weight(orc_sys* s, const orc_gc* gc, const double* b1, const double* b2, int ex_lo,
                                int ex_hi) {
  const int ch = gc->bead_charge != 0;
  static const double none[3] = {0, 0, 0};
  const double e = orc_cbmc_beads_energy(s, b1, gc->type_c, gc->bead_charge, ch ? b2 : b1, gc->type_c,
                                         -gc->bead_charge, ch, none, none, 0, ex_lo, ex_hi);
  return exp(-s->p.beta * e);
}

/* Two-stage insertion (no reference counterpart; the engine's
   plum_sim_gc_prescreen mode 2): first bead and counterion drawn as the
   growth draws them; a first stage min(1, r_1), r_1 = e^{beta mu} C w_1
   theta; on passing the rest of the growth and a second stage
   min(1, W / (w_1 theta)).  *stage: 0 rejected at the first stage, 1 past
   it.  Returns 1 when accepted. */
int orc_gc_insertion2(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double log_theta,
                      double* chain_out, double* weight_out, int* stage) {
  const int L = gc->chain_len, ch = gc->bead_charge != 0;
  memset(chain_out, 0, sizeof(double) * 6 * L);
  *stage = 0;
  if (weight_out) *weight_out = 0;
  for (int i = 0; i < (ch ? 2 : 1); i++)
    for (int j = 0; j < 3; j++) chain_out[3 * (i * L) + j] = u01(r) * s->p.box[j];
  const double w1 = first_bead_weight(s, gc, chain_out, chain_out + 3 * L, 0, -1);
  if (w1 <= 0) return 0;
  const double C = insertion_factor(s, gc);
  const double r1 = exp(s->p.beta * gc->chem_pot) * C * w1 * exp(log_theta);
  if (!(u01(r) < r1)) return 0;
  *stage = 1;
  const double weight = grow_chain(s, r, gs, gc, chain_out, 0, 1);
  if (weight_out) *weight_out = weight;
  if (weight <= 0) return 0;
  return u01(r) < weight / (w1 * exp(log_theta));
}

/* The chain to delete, chosen by the reference's draw (cbmc.cc:361-372), or
   -1 (nothing to delete, or not the exchange layout). */
static int deletion_pick(orc_sys* s, orc_mt19937* r, const orc_gc* gc) {
  const int L = gc->chain_len, ch = gc->bead_charge != 0;
  int n_chain, n_cion, n_aion;
  gc_counts(s, gc, &n_chain, &n_cion, &n_aion);
  const int spc1 = gc_spc1(gc, n_chain, n_cion, n_aion);
  int del = (int)floor(u01(r) * spc1);
  if (del >= spc1) del = spc1 - 1;
  /* nothing to delete: the reference would index molecule -(1 + L) */
  if (del < 0) return -1;
  del = gc->phantom + gc->coion + gc->grafted + gc->grafted_counterion + del * (ch ? 1 + L : 1);
  const int counterion = ch ? L : 0;
  if (del + counterion >= s->n_mols || s->mol_start[del + 1] - s->mol_start[del] != L) return -1;
  return del;
}

/* The Rosenbluth weight of chain `del` regrown along itself (cbmc.cc:373-400):
   trial 0 of each step is the existing bead. */
static double deletion_weight(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, int del) {
  const int K = gc->trials, L = gc->chain_len, ch = gc->bead_charge != 0;
  const int counterion = ch ? L : 0;
  const int second = ch ? del + 1 : del;
  double* chain = (double*)calloc(6 * (size_t)L, sizeof(double));
  double* tri = (double*)malloc(sizeof(double) * 6 * K);
  double* wts = (double*)malloc(sizeof(double) * K);
  const double* b0 = &s->cur[3 * s->mol_start[del]];
  const double* i0 = &s->cur[3 * s->mol_start[second]];
  double weight = 1.0;
  weight *= exp(-s->p.beta * orc_cbmc_beads_energy(s, b0, gc->type_c, gc->bead_charge, i0, gc->type_c,
                                                   -gc->bead_charge, ch, chain, chain + 3 * L, 0, del,
                                                   del + counterion));
  memcpy(chain, b0, sizeof(double) * 3);
  if (ch) memcpy(&chain[3 * L], i0, sizeof(double) * 3);
  for (int i = 1; i < L; i++) {
    memcpy(&chain[3 * i], &s->cur[3 * (s->mol_start[del] + i)], sizeof(double) * 3);
    if (ch) memcpy(&chain[3 * (i + L)], &s->cur[3 * s->mol_start[del + i + 1]], sizeof(double) * 3);
    gen_trials(s, r, gs, gc, &chain[3 * (i - 1)], chain, i, del, del + counterion, tri, wts);
    wts[0] = exp(-s->p.beta * orc_cbmc_beads_energy(s, &chain[3 * i], gc->type_c, gc->bead_charge,
                                                    &chain[3 * (i + (ch ? L : 0))], gc->type_c,
                                                    -gc->bead_charge, ch, chain, chain + 3 * L, i, del,
                                                    del + counterion));
    double Wi = 0;
    for (int j = 0; j < K; j++) Wi += wts[j];
    weight *= Wi / K;
  }
  free(chain);
  free(tri);
  free(wts);
  return weight;
}

int orc_gc_deletion(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double* weight_out) {
  const int del = deletion_pick(s, r, gc);
  if (del < 0) return -1;
  const double C = deletion_factor(s, gc);
  const double weight = deletion_weight(s, r, gs, gc, del);
  if (weight_out) *weight_out = weight;
  const double rn = u01(r);
  return rn < C / (exp(s->p.beta * gc->chem_pot) * weight) ? del : -1;
}

/* Two-stage deletion (the inverse of orc_gc_insertion2): the chain chosen as
   the reference chooses it, a first stage min(1, C / (e^{beta mu} w_1
   theta)) on its own first bead and counterion against the rest of the
   system, then the regrowth and a second stage min(1, w_1 theta / W).
   *stage: -1 no chain to delete, 0 rejected at the first stage, 1 past it. */
int orc_gc_deletion2(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, double log_theta,
                     double* weight_out, int* stage) {
  const int L = gc->chain_len, ch = gc->bead_charge != 0;
  *stage = -1;   /* no chain to screen */
  if (weight_out) *weight_out = 0;
  const int del = deletion_pick(s, r, gc);
  if (del < 0) return -1;
  *stage = 0;
  const int second = ch ? del + 1 : del;
  const double w1 = first_bead_weight(s, gc, &s->cur[3 * s->mol_start[del]], &s->cur[3 * s->mol_start[second]],
                                      del, del + (ch ? L : 0));
  const double C = deletion_factor(s, gc);
  const double r1 = C / (exp(s->p.beta * gc->chem_pot) * w1 * exp(log_theta));
  if (!(u01(r) < r1)) return -1;
  *stage = 1;
  const double weight = deletion_weight(s, r, gs, gc, del);
  if (weight_out) *weight_out = weight;
  return u01(r) < w1 * exp(log_theta) / weight ? del : -1;
}

double orc_gc_widom(orc_sys* s, orc_mt19937* r, orc_gauss* gs, const orc_gc* gc, int n_ins) {
  const int L = gc->chain_len;
  int n_chain, n_cion, n_aion;
  gc_counts(s, gc, &n_chain, &n_cion, &n_aion);
  const int spc1 = gc_spc1(gc, n_chain, n_cion, n_aion), spc2 = gc_spc2(gc, n_cion, n_aion);
  const int spc1_add = 1, spc2_add = gc->bead_charge == 0 ? 0 : L;
  double factorial = 1;
  for (int i = spc1 + 1; i <= spc1 + spc1_add; i++) factorial *= i;
  for (int i = spc2 + 1; i <= spc2 + spc2_add; i++) factorial *= i;
  const double vol = s->p.box[0] * s->p.box[1] * s->p.box[2];
  double vol_over_lam = 1;
  vol_over_lam *= pow(vol / 15.625 / (gc->debroglie / pow((double)L, 1.5)), spc1_add);
  vol_over_lam *= pow(vol / 15.625 / (gc->debroglie / pow(1.0, 1.5)), spc2_add);
  const double C = vol_over_lam * (1.0 / factorial);
  double* chain = (double*)calloc(6 * (size_t)L, sizeof(double));
  double total = 0;
  for (int c = 0; c < n_ins; c++) total += grow_chain(s, r, gs, gc, chain, 1, 0) * C;
  free(chain);
  return total / (double)n_ins;
}

/* ------------------------------------------------------------------ */
/* volume-perturbation pressure (pressure.cc:187-387)                   */
/* ------------------------------------------------------------------ */
static const double kDz = 0.00001; /* utilities/constants.h:55 */

/* PotentialEwaldCoul::PairEnergyRealForP (potential_ewald_coul.cc:167-196):
   the real-space sum in the box stretched along z by kDz (no early return
   for an uncharged pair: the prefactor is then 0). */
static double real_pair_forp(const orc_sys* s, const double* a, double qa, const double* b,
                             double qb) {
  double energy = 0, dist[3];
  const double* L = s->e.box_e;
  const double Ls[3] = {L[0], L[1], L[2] + kDz};
  get_dist_vector(a, b, Ls, s->p.npbc, dist);
  const double prefactor = s->p.lB * qa * qb;
  const int* c = s->e.real_cell;
  const double sa = sqrt(s->p.alpha);
  for (int i = -c[0]; i <= c[0]; i++)
    for (int j = -c[1]; j <= c[1]; j++)
      for (int k = -c[2]; k <= c[2]; k++) {
        double rv0 = dist[0] + i * L[0];
        double rv1 = dist[1] + j * L[1];
        double rv2 = dist[2] + k * (L[2] + kDz);
        double r = sqrt(rv0 * rv0 + rv1 * rv1 + rv2 * rv2);
        if (r > 0 && r <= s->e.real_cutoff) energy += prefactor * erfc(sa * r) / r;
      }
  return energy;
}

/* PotentialEwaldCoul::PairEnergyReplForP (potential_ewald_coul.cc:228-250)
   with the kz_forP / k2_forP / ek2_forP / box_vol_forP tables of
   :52-54, :86-113 (the same l grid, z wavenumbers of the stretched box). */
static double repl_pair_forp(const orc_sys* s, const double* a, double qa, const double* b,
                             double qb) {
  double energy = 0, r[3];
  const double* L = s->e.box_e;
  const double Ls[3] = {L[0], L[1], L[2] + kDz};
  get_dist_vector(a, b, Ls, s->p.npbc, r);
  const double vol_forp = L[0] * L[1] * (L[2] + kDz);
  const double prefactor = s->p.lB * qa * qb / (kPi * vol_forp) * (4 * kPi * kPi);
  const int* rc = s->e.repl_cell;
  for (int lx = 0; lx < s->ceto[0]; lx++)
    for (int ly = 0; ly < s->ceto[1]; ly++)
      for (int lz = 0; lz < s->ceto[2]; lz++) {
        const double kzp = (double)((lz - rc[2]) * 2) * kPi / (L[2] + kDz);
        const double k2p = s->kx[lx] * s->kx[lx] + s->ky[ly] * s->ky[ly] + kzp * kzp;
        if (k2p > 0 && k2p <= s->e.repl_cutoff) {
          const double ek2p = exp(-k2p / (4 * s->p.alpha)) / k2p;
          energy += prefactor * ek2p * cos(s->kx[lx] * r[0] + s->ky[ly] * r[1] + kzp * r[2]);
        }
      }
  return energy;
}

/* PotentialTruncatedLJ::PairEnergy with an explicit box (the pressure
   estimator passes the stretched box). */
static double lj_pair_box(const orc_sys* s, const double* a, int ta, const double* b, int tb,
                          const double box[3]) {
  orc_sys t = *s;
  for (int k = 0; k < 3; k++) t.p.box[k] = box[k];
  return lj_pair_xyz(&t, a, ta, b, tb);
}

/* PotentialTruncatedLJWall::BeadEnergy with the wall at z = Lz. */
static double wall_bead_zL(const orc_sys* s, double z, int t, double Lz) {
  orc_sys u = *s;
  u.p.box[2] = Lz;
  return wall_bead_z(&u, z, t);
}

/* ForceField::CalcPressureVolScalingHSELSlit (pressure.cc:187-387), the
   energy change of one stretch of the box along z by kDz: every bead's z
   scaled by 1 + kDz/Lz with the bond potential on (grafted "L" beads stay,
   "R" beads move by kDz), each molecule's z shifted by kDz com_z/Lz without
   it; all pairs once (self pairs halved in the Ewald part, phantom surface
   sites only across the two halves of the site list, LJ between non-phantom
   beads only), walls, springs and the slab dipole.
   out: [0] dU, [1] Ewald pairs, [2] LJ pairs, [3] wall, [4] springs,
   [5] dipole, [6] molecules n_mol - phantom, [7] volume Lx Ly Lz,
   [8] area x dz. */
void orc_vp_sample(orc_sys* s, int phantom, double out[9]) {
  const double* box = s->p.box;
  const int n_mol = s->n_mols;
  double* trial = s->trial;
  double* dcz = malloc(sizeof(double) * (n_mol ? n_mol : 1));
  for (int i = 0; i < n_mol; i++) {
    const int a = s->mol_start[i], b = s->mol_start[i + 1];
    double com_z = 0;
    for (int j = a; j < b; j++) com_z += s->cur[3 * j + 2];
    com_z /= (b - a);
    dcz[i] = kDz * (com_z / box[2]);
    if (s->p.type_kind[s->type[a]] == 2) dcz[i] = kDz;
    else if (s->p.type_kind[s->type[a]] == 1) dcz[i] = 0;
  }
  for (int i = 0; i < n_mol; i++)
    for (int j = s->mol_start[i]; j < s->mol_start[i + 1]; j++) {
      trial[3 * j] = s->cur[3 * j];
      trial[3 * j + 1] = s->cur[3 * j + 1];
      const int kind = s->p.type_kind[s->type[j]];
      if (s->p.use_bond && kind == 0) trial[3 * j + 2] = s->cur[3 * j + 2] * (1.0 + kDz / box[2]);
      else trial[3 * j + 2] = s->cur[3 * j + 2] + dcz[i];
    }
  const double box_sc[3] = {box[0], box[1], box[2] + kDz};
  double dU = 0, d_el = 0, d_hs = 0, d_ext = 0, d_bond = 0, d_dip = 0;
  for (int i = 0; i < n_mol; i++) {
    for (int j = s->mol_start[i]; j < s->mol_start[i + 1]; j++) {
      for (int k = i; k < n_mol; k++) {
        for (int l = s->mol_start[k]; l < s->mol_start[k + 1]; l++) {
          if (!((k > i || (k == i && l >= j)) &&
                (i >= phantom || (i < phantom / 2 && k >= phantom / 2))))
            continue;
          if (s->p.use_ewald) {
            double oldE = real_pair_xyz(s, &s->cur[3 * j], s->q[j], &s->cur[3 * l], s->q[l]) +
                          repl_pair_xyz(s, &s->cur[3 * j], s->q[j], &s->cur[3 * l], s->q[l]);
            if (j == l) oldE *= 0.5;
            double newE = real_pair_forp(s, &trial[3 * j], s->q[j], &trial[3 * l], s->q[l]);
            newE += repl_pair_forp(s, &trial[3 * j], s->q[j], &trial[3 * l], s->q[l]);
            if (j == l) newE *= 0.5;
            dU += newE - oldE;
            d_el += newE - oldE;
          }
          if (s->p.use_pair && i >= phantom && k >= phantom && j != l) {
            double oldE = lj_pair_xyz(s, &s->cur[3 * j], s->type[j], &s->cur[3 * l], s->type[l]);
            double newE = lj_pair_box(s, &trial[3 * j], s->type[j], &trial[3 * l], s->type[l], box_sc);
            dU += newE - oldE;
            d_hs += newE - oldE;
          }
        }
      }
      if (s->p.use_ext && i >= phantom) {
        double oldE = wall_bead_z(s, s->cur[3 * j + 2], s->type[j]);
        double newE = wall_bead_zL(s, trial[3 * j + 2], s->type[j], box_sc[2]);
        dU += newE - oldE;
        d_ext += newE - oldE;
      }
    }
    if (s->p.use_bond) {
      const double oldE = s->bond_cur[i];
      const double newE = spring_molecule(s, i);   /* on the trial coordinates */
      dU += newE - oldE;
      d_bond += newE - oldE;
    }
  }
  if (s->p.use_ewald && s->p.dipole) {
    double mo = 0, mn = 0;
    for (int j = 0; j < s->n; j++) {
      mo += s->q[j] * s->cur[3 * j + 2];
      mn += s->q[j] * trial[3 * j + 2];
    }
    const double vol = box[0] * box[1] * box[2];
    d_dip = (s->p.lB * 2 * kPi) * (mn * mn / (box[0] * box[1] * (box[2] + kDz)) - mo * mo / vol);
    dU += d_dip;
  }
  /* unscale: trial back to current */
  memcpy(trial, s->cur, sizeof(double) * 3 * s->n);
  free(dcz);
  out[0] = dU; out[1] = d_el; out[2] = d_hs; out[3] = d_ext; out[4] = d_bond; out[5] = d_dip;
  out[6] = n_mol - phantom;
  out[7] = box[0] * box[1] * box[2];
  out[8] = box[0] * box[1] * kDz;
}

/* ------------------------------------------------------------------ */
/* wall-force slit pressure                                            */
/* ------------------------------------------------------------------ */

/* PotentialTruncatedLJ::PairForce (potential_truncated_lj.cc:87-121); r6
   starts at 0 as in lj_pair_xyz (the reference leaves it unset when r > 0).
   Beyond a full LJ cutoff it keeps the reference's expression with
   r6 = (sigma / lj_cutoff)^6. */
static double lj_force_xyz(const orc_sys* s, const double* a, int ta, const double* b, int tb) {
  double force = 0;
  double r = bb_dist(a, b, s->p.box, s->p.npbc);
  double sigma = (s->p.lj_sigma[ta] + s->p.lj_sigma[tb]) / 2;
  double epsilon = sqrt(s->p.lj_eps[ta] * s->p.lj_eps[tb]);
  double r6 = 0;
  if (r <= 0) r6 = -1;
  if (r6 == -1) {
    force = kVeryLargeEnergy;
  } else if (s->p.lj_cutoff < 0) {
    if (r < k216 * sigma) {
      r6 = pow(sigma / r, 6);
      force = 4 * epsilon * (r6 * r6 * 12 / r - r6 * 6 / r);
    }
  } else if (r < s->p.lj_cutoff) {
    r6 = pow(sigma / r, 6);
    force = 4 * epsilon * (r6 * r6 * 12 / r - r6 * 6 / r);
  } else {
    r6 = pow(sigma / s->p.lj_cutoff, 6);
    force = 4 * epsilon * (r6 * r6 * 12 / r - r6 * 6 / r);
  }
  return force;
}

/* PotentialEwaldCoul::PairForceZReal (potential_ewald_coul.cc:261-293): z
   force on the wall site a from bead b, over the real-space images. */
static double real_force_z_xyz(const orc_sys* s, const double* a, double qa, const double* b, double qb) {
  double force_z = 0;
  if (qa * qb == 0) return force_z;
  double r[3];
  get_dist_vector(b, a, s->e.box_e, s->p.npbc, r);
  double prefactor = s->p.lB * qa * qb;
  double prefactor2 = 2 * sqrt(s->p.alpha / kPi);
  const int* c = s->e.real_cell;
  const double* L = s->e.box_e;
  for (int i = -c[0]; i <= c[0]; i++)
    for (int j = -c[1]; j <= c[1]; j++)
      for (int k = -c[2]; k <= c[2]; k++) {
        double rv0 = r[0] + i * L[0];
        double rv1 = r[1] + j * L[1];
        double rv2 = r[2] + k * L[2];
        double d = sqrt(rv0 * rv0 + rv1 * rv1 + rv2 * rv2);
        if (d > 0 && d <= s->e.real_cutoff)
          force_z += prefactor * ((prefactor2 * exp(-s->p.alpha * d * d)) + (erfc(sqrt(s->p.alpha) * d) / d)) *
                     rv2 / (d * d);
      }
  return force_z;
}

/* PotentialEwaldCoul::PairForceZRepl (potential_ewald_coul.cc:297-324). */
static double repl_force_z_xyz(const orc_sys* s, const double* a, double qa, const double* b, double qb) {
  double force_z = 0;
  if (qa * qb == 0) return force_z;
  double r[3];
  get_dist_vector(b, a, s->e.box_e, s->p.npbc, r);
  double prefactor = s->p.lB * 4 * kPi * qa * qb / s->e.vol_e;
  for (int lx = 0; lx < s->ceto[0]; lx++)
    for (int ly = 0; ly < s->ceto[1]; ly++)
      for (int lz = 0; lz < s->ceto[2]; lz++) {
        int idx = s->ceto[1] * s->ceto[2] * lx + s->ceto[2] * ly + lz;
        if (s->k2[idx] > 0 && s->k2[idx] <= s->e.repl_cutoff)
          force_z += prefactor * s->kz[lz] * s->ek2[idx] *
                     sin(s->kx[lx] * r[0] + s->ky[ly] * r[1] + s->kz[lz] * r[2]);
      }
  return force_z;
}

/* PotentialTruncatedLJWall::BeadForceOnWall (potential_truncated_lj_wall.cc:
   136-214) at z (the coordinates are settled when the pressure is sampled),
   branch for branch; the full-LJ branch keeps the reference's /z in its last
   term. */
static double wall_force_z(const orc_sys* s, double z, int t) {
  double force = 0;
  const double Lz = s->p.box[2];
  double sigma = s->p.wall_sigma[t];
  double epsilon = s->p.wall_eps[t];
  int kind = s->p.type_kind[t];
  double R0 = 3 * k213 * sigma;
  double K = 1;
  double m_cut = s->p.wall_cut;
  if (epsilon == 0) {
    force = 0;
  } else if (z <= 0 || z >= Lz) {
    return kVeryLargeEnergy;
  } else {
    if (m_cut < 0 && kind != ORC_KIND_R && kind != ORC_KIND_L) {
      if (z < k213 * sigma) {
        double r3 = pow(sigma / z, 3);
        force += 2.59807621135 * epsilon * (r3 * r3 * 6 / z - r3 * 3 / z);
      }
      if (Lz - z < k213 * sigma) {
        double r3 = pow(sigma / (Lz - z), 3);
        force += 2.59807621135 * epsilon * (r3 * r3 * 6 / (Lz - z) - r3 * 3 / (Lz - z));
      }
    } else if (kind == ORC_KIND_L) {
      double r3 = pow(sigma / z, 3);
      force += 2.59807621135 * epsilon * (r3 * r3 * 6 / z - r3 * 3 / z);
      force += -K * R0 * z / (1 - pow(z / R0, 2));
      if (Lz - z < k213 * sigma) {
        double r3b = pow(sigma / (Lz - z), 3);
        force += 2.59807621135 * epsilon * (r3b * r3b * 6 / (Lz - z) - r3b * 3 / (Lz - z));
      }
    } else if (kind == ORC_KIND_R) {
      double r3 = pow(sigma / (Lz - z), 3);
      force += 2.59807621135 * epsilon * (r3 * r3 * 6 / (Lz - z) - r3 * 3 / (Lz - z));
      force += -K * R0 * (Lz - z) / (1 - pow((Lz - z) / R0, 2));
      if (z < k213 * sigma) {
        double r3b = pow(sigma / z, 3);
        force += 2.59807621135 * epsilon * (r3b * r3b * 6 / z - r3b * 3 / z);
      }
    } else {
      if (z < m_cut) {
        double r3 = pow(sigma / z, 3);
        force += 2.59807621135 * epsilon * (r3 * r3 * 6 / z - r3 * 3 / z);
      } else {
        double r3 = pow(sigma / m_cut, 3);
        force += 2.59807621135 * epsilon * (r3 * r3 * 6 / z - r3 * 3 / z);
      }
      if (Lz - z < m_cut) {
        double r3 = pow(sigma / (Lz - z), 3);
        force += 2.59807621135 * epsilon * (r3 * r3 * 6 / (Lz - z) - r3 * 3 / (Lz - z));
      } else {
        double r3 = pow(sigma / m_cut, 3);
        force += 2.59807621135 * epsilon * (r3 * r3 * 6 / z - r3 * 3 / z);
      }
    }
  }
  return force;
}

double orc_lj_force(orc_sys* s, int a, int b) {
  return lj_force_xyz(s, &s->cur[3 * a], s->type[a], &s->cur[3 * b], s->type[b]);
}
double orc_real_force_z(orc_sys* s, int a, int b) {
  return real_force_z_xyz(s, &s->cur[3 * a], s->q[a], &s->cur[3 * b], s->q[b]);
}
double orc_repl_force_z(orc_sys* s, int a, int b) {
  return repl_force_z_xyz(s, &s->cur[3 * a], s->q[a], &s->cur[3 * b], s->q[b]);
}
double orc_wall_force(orc_sys* s, int a) { return wall_force_z(s, s->cur[3 * a + 2], s->type[a]); }

/* ForceField::CalcPressureForceLJELSlit (pressure.cc:404-470), the sums of
   one call: the loops, coefficients and classes as the reference has them. */
void orc_slit_forces(orc_sys* s, int phantom, double f[6]) {
  for (int k = 0; k < 6; k++) f[k] = 0;
  for (int i = 0; i < phantom; i++) {
    const int wi = s->mol_start[i];
    const double* pw = &s->cur[3 * wi];
    for (int j = phantom / 2; j < s->n_mols; j++) {
      const int jlen = s->mol_start[j + 1] - s->mol_start[j];
      for (int k = s->mol_start[j]; k < s->mol_start[j + 1]; k++) {
        double C = -1;
        if (i < phantom / 2 && j >= phantom) C = -0.5;
        else if (i < phantom && j >= phantom) C = 0.5;
        else C = -1;
        int k_id = 2;
        if (j >= phantom) k_id = jlen > 1 ? 1 : 0;
        const double* pb = &s->cur[3 * k];
        if (s->p.use_pair && (i < phantom / 2 || j >= phantom)) {
          double r[3];
          get_dist_vector(pb, pw, s->p.box, s->p.npbc, r);
          double z_component = r[2] / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
          f[k_id] += C * z_component * lj_force_xyz(s, pb, s->type[k], pw, s->type[wi]);
        }
        if (s->p.use_ewald && (i < phantom / 2 || j >= phantom)) {
          double force = C * real_force_z_xyz(s, pw, s->q[wi], pb, s->q[k]);
          force += C * repl_force_z_xyz(s, pw, s->q[wi], pb, s->q[k]);
          f[3 + k_id] += force;
        }
      }
    }
  }
  if (s->p.use_ext)
    for (int i = phantom; i < s->n_mols; i++) {
      const int ilen = s->mol_start[i + 1] - s->mol_start[i];
      for (int j = s->mol_start[i]; j < s->mol_start[i + 1]; j++)
        f[ilen > 1 ? 1 : 0] += 0.5 * wall_force_z(s, s->cur[3 * j + 2], s->type[j]);
    }
}
