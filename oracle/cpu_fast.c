/* CPU twin of the engine's trial-move energy algorithm, for the CPU
 * baseline only (test infrastructure: nothing in the product calls it).
 *
 * Same algorithm as the GPU path, not the reference's: LJ + real space by a
 * cell list over the Ewald box (every image of a partner within the search
 * radius found by scanning the image cells, as the engine's queries do), and
 * the k space from a running structure factor S(k) over the half-space k
 * list, dE_k = sum_k c_k (2 Re(S* dS) + |dS|^2) with dS(k) the moved beads'
 * change of S (phase tables e^{i l theta} per dimension, products of three
 * entries).  FP64 throughout, OpenMP over the moved positions (real space)
 * and over k vectors (k space).  The terms it evaluates (LJ, real, k space)
 * are checked against the oracle's EnergyDifference in
 * tests/test_cpu_fast.py; the O(moved beads) terms (springs, walls, self,
 * dipole) are left out of the twin: they are under 1% of a move's work.
 *
 * Physics conventions followed (same as the oracle's): LJ on the minimum
 * image in the simulation box (periodic dimensions only), WCA cut at
 * 2^(1/6) sigma or a shifted cut LJ (potential_truncated_lj.cc:49-85);
 * real space lB qi qj erfc(sqrt(alpha) r) / r over all images within the
 * real cutoff (potential_ewald_coul.cc:134-164); k vectors l 2 kPi / L_e with
 * the reference's kPi, c_k = lB 4 kPi^2 / (kPi V_e) e^{-k^2/(4 alpha)} / k^2
 * over the half space (k and -k paired).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "plum_oracle.h"

static const double cfPi = 3.14159265359;   /* the reference's kPi */
static const double cf216 = 1.12246204831;

typedef struct {
  int n, cap;
  int* j;
} cf_cell;

typedef struct cf_sys cf_sys;

struct cf_sys {
  orc_params p;
  orc_ewald_info e;
  int n;
  double *x, *q;          /* unwrapped positions (3n), charges */
  double* w;              /* wrapped into the Ewald box (3n) */
  int* type;
  int* cell_of;           /* cell of each bead */
  unsigned char* moved;
  /* LJ table */
  int nt;
  double *lj_s2, *lj_e4, *lj_cut, *lj_shift;
  double rs, rs2, rc, sa;
  /* cells */
  int nc[3], ncell;
  double h[3];
  cf_cell* cells;
  /* k space */
  int nk, nkd[3];         /* half-space k vectors; table extents |l| <= nkd */
  int* kl;                /* 3 per k */
  double *ck, *S, *dS;    /* S, dS: 2 per k */
  /* pending move */
  int nm;
  int* idx;
  double* xn;             /* new positions (3 nm) */
};

static double wrap1(double x, double L) {
  double y = x - L * floor(x / L);
  if (y >= L) y -= L;
  if (y < 0) y += L;
  return y;
}

static int cell_coord(double w, double h, int n) {
  int c = (int)floor(w / h);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

static int cell_of_w(const cf_sys* s, const double* w) {
  int c = 0;
  for (int k = 0; k < 3; k++) c = c * s->nc[k] + cell_coord(w[k], s->h[k], s->nc[k]);
  return c;
}

static void cell_add(cf_cell* c, int j) {
  if (c->n == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 16;
    c->j = realloc(c->j, sizeof(int) * c->cap);
  }
  c->j[c->n++] = j;
}

static void cell_remove(cf_cell* c, int j) {
  for (int i = 0; i < c->n; i++)
    if (c->j[i] == j) {
      c->j[i] = c->j[--c->n];
      return;
    }
}

static double lj_pair(const cf_sys* s, double r2, int ti, int tj) {
  const int t = ti * s->nt + tj;
  const double r = sqrt(r2);
  if (r <= 0) return 1e8;
  if (!(r < s->lj_cut[t])) return 0.0;
  const double x = s->lj_s2[t] / r2;
  const double r6 = x * x * x;
  return s->lj_e4[t] * (r6 * r6 - r6) - s->lj_shift[t];
}

/* phase tables of one position: e^{i l theta_d}, l = 0..nkd[d] per dimension */
static void tables(const cf_sys* s, const double* pos, double* t) {
  int o = 0;
  for (int d = 0; d < 3; d++) {
    const double wv = wrap1(pos[d], s->e.box_e[d]);
    for (int l = 0; l <= s->nkd[d]; l++) {
      const double th = ((double)(l * 2) * cfPi / s->e.box_e[d]) * wv;
      t[2 * o] = cos(th);
      t[2 * o + 1] = sin(th);
      o++;
    }
  }
}

static void phase(const cf_sys* s, const double* t, const int* l, double* re, double* im) {
  const int ox = l[0], oy = s->nkd[0] + 1 + abs(l[1]), oz = s->nkd[0] + 1 + s->nkd[1] + 1 + abs(l[2]);
  const double ax = t[2 * ox], bx = t[2 * ox + 1];
  const double ay = t[2 * oy], by = l[1] < 0 ? -t[2 * oy + 1] : t[2 * oy + 1];
  const double az = t[2 * oz], bz = l[2] < 0 ? -t[2 * oz + 1] : t[2 * oz + 1];
  const double xr = ax * ay - bx * by, xi = ax * by + bx * ay;
  *re = xr * az - xi * bz;
  *im = xr * bz + xi * az;
}

cf_sys* cf_create(const orc_params* p, int n, const double* xyz, const double* q, const int* type, int div) {
  cf_sys* s = calloc(1, sizeof(cf_sys));
  s->p = *p;
  s->n = n;
  orc_derive_cutoffs(p->box, p->lB, p->alpha, p->dipole, &s->e);
  s->x = malloc(sizeof(double) * 3 * n);
  s->w = malloc(sizeof(double) * 3 * n);
  s->q = malloc(sizeof(double) * n);
  s->type = malloc(sizeof(int) * n);
  s->cell_of = malloc(sizeof(int) * n);
  s->moved = calloc(n, 1);
  memcpy(s->x, xyz, sizeof(double) * 3 * n);
  memcpy(s->q, q, sizeof(double) * n);
  memcpy(s->type, type, sizeof(int) * n);
  /* LJ table */
  s->nt = p->n_types;
  const int nt2 = s->nt * s->nt;
  s->lj_s2 = malloc(sizeof(double) * nt2);
  s->lj_e4 = malloc(sizeof(double) * nt2);
  s->lj_cut = malloc(sizeof(double) * nt2);
  s->lj_shift = malloc(sizeof(double) * nt2);
  double lj_max = 0;
  for (int a = 0; a < s->nt; a++)
    for (int b = 0; b < s->nt; b++) {
      const int t = a * s->nt + b;
      const double sig = (p->lj_sigma[a] + p->lj_sigma[b]) / 2, eps = sqrt(p->lj_eps[a] * p->lj_eps[b]);
      s->lj_s2[t] = sig * sig;
      s->lj_e4[t] = 4 * eps;
      double cut, rr;
      if (p->lj_cutoff < 0) {
        cut = cf216 * sig;
        rr = 1.0 / cf216;
      } else {
        cut = p->lj_cutoff;
        rr = sig / p->lj_cutoff;
      }
      const double r6 = rr * rr * rr * rr * rr * rr;
      s->lj_cut[t] = p->use_pair ? cut : 0.0;
      s->lj_shift[t] = 4 * eps * (r6 * r6 - r6);
      if (p->use_pair && cut > lj_max) lj_max = cut;
    }
  s->rc = p->use_ewald ? s->e.real_cutoff : 0.0;
  s->sa = sqrt(p->alpha);
  s->rs = s->rc > lj_max ? s->rc : lj_max;
  if (s->rs <= 0) s->rs = 1;
  s->rs2 = s->rs * s->rs;
  /* cells over the Ewald box, edge ~ rs / div */
  s->ncell = 1;
  for (int k = 0; k < 3; k++) {
    int nc = (int)floor(s->e.box_e[k] / (s->rs / (div > 0 ? div : 2)));
    if (nc < 1) nc = 1;
    if (nc > 256) nc = 256;
    s->nc[k] = nc;
    s->h[k] = s->e.box_e[k] / nc;
    s->ncell *= nc;
  }
  s->cells = calloc(s->ncell, sizeof(cf_cell));
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) s->w[3 * i + k] = wrap1(s->x[3 * i + k], s->e.box_e[k]);
    s->cell_of[i] = cell_of_w(s, &s->w[3 * i]);
    cell_add(&s->cells[s->cell_of[i]], i);
  }
  /* half-space k list and S(k) */
  if (p->use_ewald) {
    for (int d = 0; d < 3; d++) s->nkd[d] = s->e.repl_cell[d] > 1 ? s->e.repl_cell[d] : 1;
    const double pref0 = p->lB / (cfPi * s->e.vol_e) * (4 * cfPi * cfPi);
    int cap = 1024, nk = 0;
    s->kl = malloc(sizeof(int) * 3 * cap);
    s->ck = malloc(sizeof(double) * cap);
    const int* rc3 = s->e.repl_cell;
    for (int lx = 0; lx <= rc3[0]; lx++)
      for (int ly = -rc3[1]; ly <= rc3[1]; ly++)
        for (int lz = -rc3[2]; lz <= rc3[2]; lz++) {
          if (!(lx > 0 || (lx == 0 && (ly > 0 || (ly == 0 && lz > 0))))) continue;
          const double kx = (double)(lx * 2) * cfPi / s->e.box_e[0];
          const double ky = (double)(ly * 2) * cfPi / s->e.box_e[1];
          const double kz = (double)(lz * 2) * cfPi / s->e.box_e[2];
          const double k2 = kx * kx + ky * ky + kz * kz;
          if (!(k2 > 0 && k2 <= s->e.repl_cutoff)) continue;
          if (nk == cap) {
            cap *= 2;
            s->kl = realloc(s->kl, sizeof(int) * 3 * cap);
            s->ck = realloc(s->ck, sizeof(double) * cap);
          }
          s->kl[3 * nk] = lx;
          s->kl[3 * nk + 1] = ly;
          s->kl[3 * nk + 2] = lz;
          s->ck[nk] = pref0 * (exp(-k2 / (4 * p->alpha)) / k2);
          nk++;
        }
    s->nk = nk;
    s->S = calloc(2 * (size_t)nk, sizeof(double));
    s->dS = calloc(2 * (size_t)nk, sizeof(double));
    const int tw = s->nkd[0] + s->nkd[1] + s->nkd[2] + 3;
#pragma omp parallel
    {
      double* t = malloc(sizeof(double) * 2 * tw);
      double* loc = calloc(2 * (size_t)nk, sizeof(double));
#pragma omp for schedule(static)
      for (int i = 0; i < n; i++) {
        if (q[i] == 0) continue;
        tables(s, &s->x[3 * i], t);
        for (int k = 0; k < nk; k++) {
          double re, im;
          phase(s, t, &s->kl[3 * k], &re, &im);
          loc[2 * k] += q[i] * re;
          loc[2 * k + 1] += q[i] * im;
        }
      }
#pragma omp critical
      for (int k = 0; k < 2 * nk; k++) s->S[k] += loc[k];
      free(loc);
      free(t);
    }
  }
  s->idx = malloc(sizeof(int) * n);
  s->xn = malloc(sizeof(double) * 3 * n);
  return s;
}

void cf_destroy(cf_sys* s) {
  if (!s) return;
  for (int c = 0; c < s->ncell; c++) free(s->cells[c].j);
  free(s->cells);
  free(s->x); free(s->w); free(s->q); free(s->type); free(s->cell_of); free(s->moved);
  free(s->lj_s2); free(s->lj_e4); free(s->lj_cut); free(s->lj_shift);
  free(s->kl); free(s->ck); free(s->S); free(s->dS);
  free(s->idx); free(s->xn);
  free(s);
}

/* LJ + real space of a point (wrapped pw, charge qi, type ti) against every
   unmoved bead image within rs. */
static void scan_point(const cf_sys* s, const double* pw, double qi, int ti, double* elj, double* ere) {
  int lo[3], nd[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)floor((pw[k] - s->rs) / s->h[k]);
    nd[k] = (int)floor((pw[k] + s->rs) / s->h[k]) - lo[k] + 1;
  }
  double a = 0, b = 0;
  for (int cx = 0; cx < nd[0]; cx++)
    for (int cy = 0; cy < nd[1]; cy++)
      for (int cz = 0; cz < nd[2]; cz++) {
        const int u[3] = {lo[0] + cx, lo[1] + cy, lo[2] + cz};
        int cell = 0;
        double sh[3];
        int iz = 0;
        for (int k = 0; k < 3; k++) {
          const int qk = (int)floor((double)u[k] / s->nc[k]);
          cell = cell * s->nc[k] + (u[k] - qk * s->nc[k]);
          sh[k] = qk * s->e.box_e[k];
          if (k == 2) iz = qk;
        }
        const int lj_ok = s->p.use_pair && (s->p.npbc == 3 || iz == 0);
        const cf_cell* c = &s->cells[cell];
        for (int t = 0; t < c->n; t++) {
          const int j = c->j[t];
          if (s->moved[j]) continue;
          const double dx = s->w[3 * j] + sh[0] - pw[0], dy = s->w[3 * j + 1] + sh[1] - pw[1],
                       dz = s->w[3 * j + 2] + sh[2] - pw[2];
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 > s->rs2) continue;
          if (lj_ok) a += lj_pair(s, r2, ti, s->type[j]);
          const double qq = qi * s->q[j];
          if (s->p.use_ewald && qq != 0) {
            const double r = sqrt(r2);
            if (r > 0 && r <= s->rc) b += s->p.lB * qq * erfc(s->sa * r) / r;
          }
        }
      }
  *elj += a;
  *ere += b;
}

/* LJ (minimum image, simulation box) + real space (all images within rc) of
   two moved positions. */
static void pair_mm(const cf_sys* s, const double* a, int ta, double qa, const double* b, int tb, double qb,
                    double* elj, double* ere) {
  double d[3];
  for (int k = 0; k < 3; k++) {
    double dk = b[k] - a[k];
    if (k < s->p.npbc) dk -= s->p.box[k] * round(dk / s->p.box[k]);
    d[k] = dk;
  }
  if (s->p.use_pair) *elj += lj_pair(s, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], ta, tb);
  if (!s->p.use_ewald || qa * qb == 0) return;
  double e[3];
  for (int k = 0; k < 3; k++) {
    double dk = b[k] - a[k];
    if (k < s->p.npbc) dk -= s->e.box_e[k] * round(dk / s->e.box_e[k]);
    e[k] = dk;
  }
  const int* rc = s->e.real_cell;
  for (int i = -rc[0]; i <= rc[0]; i++)
    for (int j = -rc[1]; j <= rc[1]; j++)
      for (int k = -rc[2]; k <= rc[2]; k++) {
        const double x = e[0] + i * s->e.box_e[0], y = e[1] + j * s->e.box_e[1], z = e[2] + k * s->e.box_e[2];
        const double r = sqrt(x * x + y * y + z * z);
        if (r > 0 && r <= s->rc) *ere += s->p.lB * qa * qb * erfc(s->sa * r) / r;
      }
}

void cf_delta(cf_sys* s, int nm, const int* idx, const double* xyz_new, double out[3]) {
  s->nm = nm;
  memcpy(s->idx, idx, sizeof(int) * nm);
  memcpy(s->xn, xyz_new, sizeof(double) * 3 * nm);
  for (int m = 0; m < nm; m++) s->moved[idx[m]] = 1;
  double elj = 0, ere = 0;
  /* moved beads against the rest, old and new positions */
#pragma omp parallel for reduction(+ : elj, ere) schedule(dynamic)
  for (int t = 0; t < 2 * nm; t++) {
    const int m = t >> 1, on = t & 1, i = idx[m];
    double pw[3], a = 0, b = 0;
    for (int k = 0; k < 3; k++) pw[k] = on ? wrap1(xyz_new[3 * m + k], s->e.box_e[k]) : s->w[3 * i + k];
    scan_point(s, pw, s->q[i], s->type[i], &a, &b);
    elj += on ? a : -a;
    ere += on ? b : -b;
  }
  /* moved-moved pairs, new minus old */
  for (int m = 0; m < nm; m++)
    for (int u = m + 1; u < nm; u++) {
      const int i = idx[m], j = idx[u];
      double a = 0, b = 0, c = 0, d = 0;
      pair_mm(s, &xyz_new[3 * m], s->type[i], s->q[i], &xyz_new[3 * u], s->type[j], s->q[j], &a, &b);
      pair_mm(s, &s->x[3 * i], s->type[i], s->q[i], &s->x[3 * j], s->type[j], s->q[j], &c, &d);
      elj += a - c;
      ere += b - d;
    }
  /* k space: dS(k) and dE_k */
  double ek = 0;
  if (s->p.use_ewald && s->nk > 0) {
    const int tw = s->nkd[0] + s->nkd[1] + s->nkd[2] + 3;
    double* tab = malloc(sizeof(double) * 2 * tw * 2 * nm);
    for (int m = 0; m < nm; m++) {
      tables(s, &s->x[3 * idx[m]], &tab[2 * tw * (2 * m)]);
      tables(s, &xyz_new[3 * m], &tab[2 * tw * (2 * m + 1)]);
    }
#pragma omp parallel for reduction(+ : ek) schedule(static)
    for (int k = 0; k < s->nk; k++) {
      double dr = 0, di = 0;
      for (int m = 0; m < nm; m++) {
        const double qm = s->q[idx[m]];
        if (qm == 0) continue;
        double r0, i0, r1, i1;
        phase(s, &tab[2 * tw * (2 * m)], &s->kl[3 * k], &r0, &i0);
        phase(s, &tab[2 * tw * (2 * m + 1)], &s->kl[3 * k], &r1, &i1);
        dr += qm * (r1 - r0);
        di += qm * (i1 - i0);
      }
      s->dS[2 * k] = dr;
      s->dS[2 * k + 1] = di;
      ek += s->ck[k] * (2.0 * (s->S[2 * k] * dr + s->S[2 * k + 1] * di) + dr * dr + di * di);
    }
    free(tab);
  }
  for (int m = 0; m < nm; m++) s->moved[idx[m]] = 0;
  out[0] = elj;
  out[1] = ere;
  out[2] = ek;
}

void cf_commit(cf_sys* s) {
  for (int m = 0; m < s->nm; m++) {
    const int i = s->idx[m];
    for (int k = 0; k < 3; k++) {
      s->x[3 * i + k] = s->xn[3 * m + k];
      s->w[3 * i + k] = wrap1(s->xn[3 * m + k], s->e.box_e[k]);
    }
    const int c = cell_of_w(s, &s->w[3 * i]);
    if (c != s->cell_of[i]) {
      cell_remove(&s->cells[s->cell_of[i]], i);
      cell_add(&s->cells[c], i);
      s->cell_of[i] = c;
    }
  }
  if (s->p.use_ewald)
    for (int k = 0; k < 2 * s->nk; k++) s->S[k] += s->dS[k];
  s->nm = 0;
}

/* The decided proposals of a run, in order: record r moves n_moved[r] beads
   idx[r * maxm ..] to xyz[3 (r * maxm) ..]; accepted ones are committed.
   out_de (3 per record, may be NULL) receives LJ, real and k-space dE.
   Returns the wall time of the loop in seconds. */
double cf_replay(cf_sys* s, int nrec, const int* n_moved, const int* idx, const double* xyz, const int* accepted,
                 int maxm, double* out_de) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < nrec; r++) {
    double d[3];
    cf_delta(s, n_moved[r], idx + (size_t)r * maxm, xyz + 3 * (size_t)r * maxm, d);
    if (out_de) memcpy(out_de + 3 * (size_t)r, d, sizeof(d));
    if (accepted[r]) cf_commit(s);
    else s->nm = 0;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

int cf_threads(void) { return omp_get_max_threads(); }
