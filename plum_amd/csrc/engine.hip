// MI355X (gfx950) trial-move energy engine: kernels and the C ABI of
// include/plum_gpu.h.
//
// Data layout in HBM (one system per context):
//   beads (index order):  x,y,z current unwrapped FP64, q FP64, type, molecule
//   cell list:            per cell a count and `cap` slots; slot-major copies
//                         of (wrapped x,y,z,q) as double4 and the bead index,
//                         so the real-space sweep streams contiguous slots
//   k space:              half-space k list (lx,ly,lz) int4, c_k FP64,
//                         running S(k) double2, pending dS(k) double2
//   pending move:         moved list (index, old/new position+charge) and the
//                         per-bead phase tables e^{i l theta_d}, l <= n_d
//   control block:        dipole moment, dipole lag state, running totals,
//                         the last dE per term, status flags
#include "engine_dev.h"
#include "plum_gpu.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace plum;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                               \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess)                                                       \
      return fail(PLUM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kMoveFlagReject = 0;

// Move record: what the caller (host path) or the device proposal writes.
struct MoveRec {
  int n_moved;
  int mol;
  int move_type;
  int valid;
};

// Device control block.
struct Ctl {
  double Mz;            // sum q z over current coordinates
  double cur_dipl;      // reference's current_dipl_E
  double trial_dipl;    // reference's trial_dipl_E
  double bond_trial;    // spring energy of the moved molecule (trial)
  double terms[8];      // dE of the pending move, per term
  double total;         // dE total with the reference's early exits
  double running[8];    // running totals per term
  double full[8];       // last from-scratch energy
  int accept;
  int overflow;         // a cell overflowed (cell list incomplete)
  int pending;
  int pad;
  unsigned long long move_counter;
  long long attempted[5];
  long long accepted[5];
};

struct Dev {
  double *x, *y, *z, *q;
  int *type, *mol_of, *mol_start;
  int *bcell, *bslot;
  unsigned char* moved;
  int* mv_slot;
  int *ccount, *cbead;
  double4* cpos;
  int* ctype;
  int4* kl;
  double* ck;
  double2 *Sk, *dS;
  double2* tab;
  LJPair* lj;
  WallType* wall;
  double* bond_cur;
  MoveRec* rec;
  int* mv_idx;
  double* mv_xyz;       // trial xyz of the move record (3 per bead)
  double4 *mv_old, *mv_new;
  int* mv_type;
  Ctl* ctl;
  double* part_k;
  double2* part_r;
  double2* part_mm;
  double2* sk_part;
  double* fpart;        // full-energy partials
};

// ============================================================ move kernels
// Phase tables and moved-list set-up for the pending move.
__global__ void k_prep(Geo g, Dev d) {
  const int nm = d.rec->n_moved;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < nm) {
    const int i = d.mv_idx[gid];
    d.moved[i] = 1;
    d.mv_slot[i] = gid;
    const double qi = d.q[i];
    d.mv_old[gid] = make_double4(d.x[i], d.y[i], d.z[i], qi);
    d.mv_new[gid] = make_double4(d.mv_xyz[3 * gid], d.mv_xyz[3 * gid + 1],
                                 d.mv_xyz[3 * gid + 2], qi);
    d.mv_type[gid] = d.type[i];
  }
  if (!g.use_ewald) return;
  const int total = nm * 2 * g.tab_w;
  for (int e = gid; e < total; e += gridDim.x * blockDim.x) {
    const int row = e / g.tab_w, off = e % g.tab_w;
    const int m = row >> 1, on = row & 1;
    const int i = d.mv_idx[m];
    int dim, l;
    if (off <= g.nkx) { dim = 0; l = off; }
    else if (off <= g.nkx + 2 * g.nky + 1) { dim = 1; l = off - (g.nkx + 1) - g.nky; }
    else { dim = 2; l = off - (g.nkx + 1) - (2 * g.nky + 1) - g.nkz; }
    double xv = on ? d.mv_xyz[3 * m + dim] : (dim == 0 ? d.x[i] : (dim == 1 ? d.y[i] : d.z[i]));
    double xw = wrap_coord(xv, g.L[dim], g.inv_L[dim]);
    // k_d = (l*2) kPi / L_d exactly as potential_ewald_coul.cc:89-96
    double kd = (double)(l * 2) * kPi / g.L[dim];
    double s, c;
    sincos(kd * xw, &s, &c);
    d.tab[(size_t)row * g.tab_w + off] = make_double2(c, s);
  }
}

// dS(k) = sum_m q_m (e^{ik.r'_m} - e^{ik.r_m}) and the k-space dE
// contribution c_k (|S+dS|^2 - |S|^2), one thread per half-space k vector.
__global__ __launch_bounds__(kBlock) void k_kspace_delta(Geo g, Dev d) {
  __shared__ double sh[kBlock / kWave];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  const int nm = d.rec->n_moved;
  double e = 0;
  if (k < g.n_kh) {
    const int4 l = d.kl[k];
    const int ox = l.x, oy = g.nkx + 1 + l.y + g.nky,
              oz = g.nkx + 1 + 2 * g.nky + 1 + l.z + g.nkz;
    double2 dS = make_double2(0, 0);
    for (int m = 0; m < nm; m++) {
      const double qm = d.mv_old[m].w;
      const double2* to = d.tab + (size_t)(2 * m) * g.tab_w;
      const double2* tn = to + g.tab_w;
      double2 eo = cmul(cmul(to[ox], to[oy]), to[oz]);
      double2 en = cmul(cmul(tn[ox], tn[oy]), tn[oz]);
      dS.x += qm * (en.x - eo.x);
      dS.y += qm * (en.y - eo.y);
    }
    const double2 S = d.Sk[k];
    e = d.ck[k] * (2.0 * (S.x * dS.x + S.y * dS.y) + dS.x * dS.x + dS.y * dS.y);
    d.dS[k] = dS;
  }
  double r = block_sum<kBlock>(e, sh);
  if (threadIdx.x == 0) d.part_k[blockIdx.x] = r;
}

// Inclusive scan of n <= 1024 ints in LDS (Hillis-Steele, blockDim = 256).
__device__ void lds_scan(int* a, int n) {
  for (int off = 1; off < n; off <<= 1) {
    int v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int i = threadIdx.x + r * kBlock;
      v[r] = (i < n && i >= off) ? a[i - off] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int i = threadIdx.x + r * kBlock;
      if (i < n) a[i] += v[r];
    }
    __syncthreads();
  }
}

// Real-space (LJ + Ewald erfc) energy of one moved bead, at its old (sign -1)
// or new (sign +1) position, against every unmoved bead image within the
// search radius.  One work-group per (moved bead, old/new).
__global__ __launch_bounds__(kBlock) void k_real_delta(Geo g, Dev d) {
  __shared__ int s_cell[kMaxQueryCells];
  __shared__ int s_img[kMaxQueryCells];
  __shared__ int s_end[kMaxQueryCells];
  __shared__ double sh[kBlock / kWave];
  const int b = blockIdx.x;
  const int m = b >> 1, on = b & 1;
  const int nm = d.rec->n_moved;
  if (m >= nm) {
    if (threadIdx.x == 0) d.part_r[b] = make_double2(0, 0);
    return;
  }
  const double4 p = on ? d.mv_new[m] : d.mv_old[m];
  const int tm = d.mv_type[m];
  const double pw[3] = {wrap_coord(p.x, g.L[0], g.inv_L[0]),
                        wrap_coord(p.y, g.L[1], g.inv_L[1]),
                        wrap_coord(p.z, g.L[2], g.inv_L[2])};
  int lo[3], nd[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)floor((pw[k] - g.rs) * g.inv_h[k]);
    int hi = (int)floor((pw[k] + g.rs) * g.inv_h[k]);
    nd[k] = hi - lo[k] + 1;
  }
  const int ncq = nd[0] * nd[1] * nd[2];
  double elj = 0, ere = 0;
  const LJPair* ljrow = d.lj + tm * g.n_types;
  for (int base = 0; base < ncq; base += kMaxQueryCells) {
    const int nloc = min(kMaxQueryCells, ncq - base);
    for (int t = threadIdx.x; t < nloc; t += kBlock) {
      int c = base + t;
      int cz = c % nd[2], cy = (c / nd[2]) % nd[1], cx = c / (nd[2] * nd[1]);
      int u[3] = {lo[0] + cx, lo[1] + cy, lo[2] + cz};
      int cell = 0, img = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        int q = floordiv(u[k], g.nc[k]);
        int w = u[k] - q * g.nc[k];
        cell = cell * g.nc[k] + w;
        img |= ((q + 128) & 0xff) << (8 * k);
      }
      s_cell[t] = cell;
      s_img[t] = img;
      s_end[t] = d.ccount[cell];
    }
    __syncthreads();
    lds_scan(s_end, nloc);
    const int total = s_end[nloc - 1];
    for (int it = threadIdx.x; it < total; it += kBlock) {
      int lo_i = 0, hi_i = nloc - 1;
      while (lo_i < hi_i) {
        int mid = (lo_i + hi_i) >> 1;
        if (s_end[mid] > it) hi_i = mid; else lo_i = mid + 1;
      }
      const int t = lo_i;
      const int slot = it - (t ? s_end[t - 1] : 0);
      const int cell = s_cell[t];
      const size_t si = (size_t)cell * g.cap + slot;
      const int j = d.cbead[si];
      if (d.moved[j]) continue;
      const double4 pj = d.cpos[si];
      const int img = s_img[t];
      const int ix = (img & 0xff) - 128, iy = ((img >> 8) & 0xff) - 128,
                iz = ((img >> 16) & 0xff) - 128;
      const double dx = pj.x + ix * g.L[0] - pw[0];
      const double dy = pj.y + iy * g.L[1] - pw[1];
      const double dz = pj.z + iz * g.L[2] - pw[2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > g.rs2) continue;
      const double r = sqrt(r2);
      if (g.use_pair && (g.npbc == 3 || iz == 0)) {
        const LJPair lp = ljrow[d.ctype[si]];
        elj += lj_energy(r, lp);
      }
      if (g.use_ewald) {
        const double qq = p.w * pj.w;
        if (qq != 0 && r > 0 && r <= g.rc) ere += g.lB * qq * erfc(g.sqrt_alpha * r) / r;
      }
    }
    __syncthreads();
  }
  const double sgn = on ? 1.0 : -1.0;
  double a = block_sum<kBlock>(elj, sh);
  double c = block_sum<kBlock>(ere, sh);
  if (threadIdx.x == 0) d.part_r[b] = make_double2(sgn * a, sgn * c);
}

// Pairs among the moved beads (all in one molecule): new-new minus old-old,
// with the reference's own pair functions.
__global__ __launch_bounds__(kBlock) void k_mm_delta(Geo g, Dev d) {
  __shared__ double sh[kBlock / kWave];
  const int nm = d.rec->n_moved;
  const long long npair = (long long)nm * (nm - 1) / 2;
  double elj = 0, ere = 0;
  for (long long pidx = blockIdx.x * (long long)kBlock + threadIdx.x; pidx < npair;
       pidx += (long long)gridDim.x * kBlock) {
    // decode pidx -> (a, b), a < b, row-major over a
    double fa = floor(((2.0 * nm - 1) - sqrt((2.0 * nm - 1) * (2.0 * nm - 1) - 8.0 * pidx)) / 2);
    long long a = (long long)fa;
    if (a < 0) a = 0;
    while (a > 0 && a * (2LL * nm - a - 1) / 2 > pidx) a--;
    while ((a + 1) * (2LL * nm - a - 2) / 2 <= pidx) a++;
    long long b = pidx - a * (2LL * nm - a - 1) / 2 + a + 1;
    const double4 on_a = d.mv_new[a], on_b = d.mv_new[b];
    const double4 oo_a = d.mv_old[a], oo_b = d.mv_old[b];
    const double na[3] = {on_a.x, on_a.y, on_a.z}, nb[3] = {on_b.x, on_b.y, on_b.z};
    const double oa[3] = {oo_a.x, oo_a.y, oo_a.z}, ob[3] = {oo_b.x, oo_b.y, oo_b.z};
    if (g.use_pair) {
      const LJPair lp = d.lj[d.mv_type[a] * g.n_types + d.mv_type[b]];
      double rn = bb_dist(na, nb, g.Lsim, g.npbc);
      double ro = bb_dist(oa, ob, g.Lsim, g.npbc);
      elj += lj_energy(rn, lp) - lj_energy(ro, lp);
    }
    if (g.use_ewald) {
      ere += real_pair_ref(na, on_a.w, nb, on_b.w, g) - real_pair_ref(oa, oo_a.w, ob, oo_b.w, g);
    }
  }
  double a = block_sum<kBlock>(elj, sh);
  double c = block_sum<kBlock>(ere, sh);
  if (threadIdx.x == 0) d.part_mm[blockIdx.x] = make_double2(a, c);
}


// Spring energy of molecule `mol` in the reference's summation order
// (potential_spring.cc:22-47, consecutive beads, no PBC).  Per-bond terms are
// evaluated in parallel into LDS and summed in bond order by thread 0, so the
// same bits come out whichever kernel evaluates the molecule.  `trial`
// selects the moved-list coordinates for moved beads.
__device__ double molecule_spring(const Geo& g, const Dev& d, int mol, bool trial,
                                  double* sbuf /* kBlock doubles */) {
  const int a = d.mol_start[mol], b = d.mol_start[mol + 1];
  double acc = 0;
  for (int base = a; base < b - 1; base += kBlock) {
    const int i = base + threadIdx.x;
    if (i < b - 1) {
      double p0[3] = {d.x[i], d.y[i], d.z[i]};
      double p1[3] = {d.x[i + 1], d.y[i + 1], d.z[i + 1]};
      if (trial) {
        const int s0 = d.mv_slot[i], s1 = d.mv_slot[i + 1];
        if (s0 >= 0) { double4 v = d.mv_new[s0]; p0[0] = v.x; p0[1] = v.y; p0[2] = v.z; }
        if (s1 >= 0) { double4 v = d.mv_new[s1]; p1[0] = v.x; p1[1] = v.y; p1[2] = v.z; }
      }
      const double dx = p0[0] - p1[0], dy = p0[1] - p1[1], dz = p0[2] - p1[2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      sbuf[threadIdx.x] = 0.5 * g.k_bond * (r - g.r0) * (r - g.r0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int cnt = min(kBlock, b - 1 - base);
      for (int k = 0; k < cnt; k++) acc += sbuf[k];
    }
    __syncthreads();
  }
  return acc;  // valid in thread 0
}

// Reduction of all partials, bonded / wall / self / dipole terms, and the
// ForceField::EnergyDifference combination with its early exits
// (force_field.cc:407-434).  With `decide` the Metropolis test
// (simulation.cc:327-332) is taken on the device from the Philox stream.
__global__ __launch_bounds__(kBlock) void k_epilogue(Geo g, Dev d, int nkb, int nrb,
                                                     int nmmb, int decide,
                                                     unsigned long long seed) {
  __shared__ double sh[kBlock / kWave];
  __shared__ double sbuf[kBlock];
  __shared__ int s_forbid;
  Ctl* C = d.ctl;
  const int nm = d.rec->n_moved;
  const int mol = d.rec->mol;
  double vk = 0, vl = 0, vr = 0;
  for (int i = threadIdx.x; i < nkb; i += kBlock) vk += d.part_k[i];
  for (int i = threadIdx.x; i < nrb; i += kBlock) { vl += d.part_r[i].x; vr += d.part_r[i].y; }
  for (int i = threadIdx.x; i < nmmb; i += kBlock) { vl += d.part_mm[i].x; vr += d.part_mm[i].y; }
  const double dk = block_sum<kBlock>(vk, sh);
  const double dl = block_sum<kBlock>(vl, sh);
  const double dr = block_sum<kBlock>(vr, sh);
  const int a = d.mol_start[mol], bend = d.mol_start[mol + 1];
  double ebt = 0;
  if (g.use_bond && bend - a > 1) ebt = molecule_spring(g, d, mol, true, sbuf);
  if (threadIdx.x == 0) s_forbid = 0;
  __syncthreads();
  double ew = 0, dmz = 0;
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    const double4 pn = d.mv_new[k], po = d.mv_old[k];
    if (g.use_ext) {
      const WallType w = d.wall[d.mv_type[k]];
      const double en = wall_energy(pn.z, w, g);
      if (en >= kVeryLargeEnergy) atomicOr(&s_forbid, 1);
      ew += en - wall_energy(po.z, w, g);
    }
    dmz += pn.w * (pn.z - po.z);
  }
  const double ewt = block_sum<kBlock>(ew, sh);
  const double dmzt = block_sum<kBlock>(dmz, sh);
  if (threadIdx.x == 0) {
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double dE = 0;
    bool early = false;
    if (g.use_pair) {
      t[T_LJ] = dl;
      dE += dl;
      if (dE >= kVeryLargeEnergy) early = true;
    }
    if (!early && g.use_ext) {
      // potential_external.cc:108-111: the first forbidden bead returns 1e8
      t[T_WALL] = s_forbid ? kVeryLargeEnergy : ewt;
      dE += t[T_WALL];
      if (dE >= kVeryLargeEnergy) early = true;
    }
    if (!early && g.use_ewald) {
      t[T_REAL] = dr;
      t[T_KSPACE] = dk;
      t[T_SELF] = 0.0;
      double dd = 0;
      if (g.dipole) {
        if (g.dipole_exact) {
          const double mz1 = C->Mz + dmzt;
          C->trial_dipl = g.dip_coef * mz1 * mz1;
          dd = C->trial_dipl - g.dip_coef * C->Mz * C->Mz;
        } else {
          // potential_ewald.cc:528-531: dipole of the CURRENT coordinates
          C->trial_dipl = g.dip_coef * C->Mz * C->Mz;
          dd = C->trial_dipl - C->cur_dipl;
        }
      }
      t[T_DIPOLE] = dd;
      dE += dr + dk + dd;
    }
    if (!early && g.use_bond && bend - a > 1) {
      t[T_BOND] = ebt - d.bond_cur[mol];
      dE += t[T_BOND];
    }
    C->bond_trial = ebt;
    for (int k = 0; k < 8; k++) C->terms[k] = t[k];
    C->total = dE;
    C->pending = 1;
    if (decide) {
      int acc = 0;
      if (d.rec->valid && dE < kVeryLargeEnergy) {
        const unsigned long long ctr = C->move_counter;
        const uint4 r = Philox::gen(make_uint4((uint32_t)ctr, (uint32_t)(ctr >> 32), 0xACCu, 7u),
                                    make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
        acc = u01_32(r.x) < exp(-g.beta * dE);
      }
      C->accept = acc;
    }
  }
}

// Commit or roll back the pending move (potential_*::FinalizeEnergyBothMaps
// and Simulation's UpdateCurrentPos / UpdateTrialPos / UnsetMoved).  The cell
// list is edited by one thread in moved-list order (deterministic).
__global__ __launch_bounds__(kBlock) void k_commit(Geo g, Dev d, int accept_host, int use_dev) {
  Ctl* C = d.ctl;
  const int nm = d.rec->n_moved;
  const int mol = d.rec->mol;
  const int acc = (use_dev ? C->accept : accept_host) && d.rec->valid;
  if (threadIdx.x == 0) {
    if (acc) {
      double dmz = 0;
      for (int k = 0; k < nm; k++) {
        const int i = d.mv_idx[k];
        const double4 pn = d.mv_new[k], po = d.mv_old[k];
        dmz += pn.w * (pn.z - po.z);
        d.x[i] = pn.x; d.y[i] = pn.y; d.z[i] = pn.z;
        const double xw = wrap_coord(pn.x, g.L[0], g.inv_L[0]);
        const double yw = wrap_coord(pn.y, g.L[1], g.inv_L[1]);
        const double zw = wrap_coord(pn.z, g.L[2], g.inv_L[2]);
        const int nc = (cell_coord(xw, g.inv_h[0], g.nc[0]) * g.nc[1] +
                        cell_coord(yw, g.inv_h[1], g.nc[1])) * g.nc[2] +
                       cell_coord(zw, g.inv_h[2], g.nc[2]);
        const int oc = d.bcell[i], os = d.bslot[i];
        if (oc == nc && oc >= 0) {
          d.cpos[(size_t)oc * g.cap + os] = make_double4(xw, yw, zw, pn.w);
          continue;
        }
        if (oc >= 0) {
          const int last = d.ccount[oc] - 1;
          if (os != last) {
            const size_t sl = (size_t)oc * g.cap + last, so = (size_t)oc * g.cap + os;
            const int jb = d.cbead[sl];
            d.cbead[so] = jb;
            d.cpos[so] = d.cpos[sl];
            d.ctype[so] = d.ctype[sl];
            d.bslot[jb] = os;
          }
          d.ccount[oc] = last;
        }
        const int ns = d.ccount[nc];
        if (ns < g.cap) {
          const size_t sn = (size_t)nc * g.cap + ns;
          d.cbead[sn] = i;
          d.cpos[sn] = make_double4(xw, yw, zw, pn.w);
          d.ctype[sn] = d.type[i];
          d.ccount[nc] = ns + 1;
          d.bcell[i] = nc;
          d.bslot[i] = ns;
        } else {
          C->overflow = 1;
          d.bcell[i] = -1;
          d.bslot[i] = -1;
        }
      }
      C->Mz += dmz;
      if (g.use_bond && d.mol_start[mol + 1] - d.mol_start[mol] > 1) d.bond_cur[mol] = C->bond_trial;
      if (g.dipole) C->cur_dipl = C->trial_dipl;
      for (int k = 0; k < 8; k++) C->running[k] += C->terms[k];
    } else {
      if (g.dipole) C->trial_dipl = C->cur_dipl;
    }
    if (d.rec->valid) {
      const int mt = d.rec->move_type;
      if (mt >= 0 && mt < 5) {
        C->attempted[mt] += 1;
        if (acc) C->accepted[mt] += 1;
      }
    }
    C->pending = 0;
    C->move_counter += 1;
  }
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    const int i = d.mv_idx[k];
    d.moved[i] = 0;
    d.mv_slot[i] = -1;
  }
}

__global__ __launch_bounds__(kBlock) void k_commit_sk(Geo g, Dev d, int accept_host, int use_dev) {
  const int acc = (use_dev ? d.ctl->accept : accept_host) && d.rec->valid;
  if (!acc) return;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < g.n_kh) {
    double2 s = d.Sk[k];
    const double2 ds = d.dS[k];
    s.x += ds.x;
    s.y += ds.y;
    d.Sk[k] = s;
  }
}

// ============================================================ from scratch
// Real space + LJ over all pairs: one thread per occupied cell slot; each
// pair (and each self image, weight 1/2) is counted from its lower index.
__global__ __launch_bounds__(kBlock) void k_full_real(Geo g, Dev d) {
  __shared__ double sh[kBlock / kWave];
  const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
  double elj = 0, ere = 0;
  if (gid < (long long)g.ncell * g.cap) {
    const int cell = (int)(gid / g.cap), slot = (int)(gid % g.cap);
    if (slot < d.ccount[cell]) {
      const size_t si = (size_t)cell * g.cap + slot;
      const int i = d.cbead[si];
      const double4 p = d.cpos[si];
      const LJPair* ljrow = d.lj + d.ctype[si] * g.n_types;
      const double pw[3] = {p.x, p.y, p.z};
      int lo[3], hi[3];
      for (int k = 0; k < 3; k++) {
        lo[k] = (int)floor((pw[k] - g.rs) * g.inv_h[k]);
        hi[k] = (int)floor((pw[k] + g.rs) * g.inv_h[k]);
      }
      for (int ux = lo[0]; ux <= hi[0]; ux++) {
        const int qx = floordiv(ux, g.nc[0]), wx = ux - qx * g.nc[0];
        for (int uy = lo[1]; uy <= hi[1]; uy++) {
          const int qy = floordiv(uy, g.nc[1]), wy = uy - qy * g.nc[1];
          for (int uz = lo[2]; uz <= hi[2]; uz++) {
            const int qz = floordiv(uz, g.nc[2]), wz = uz - qz * g.nc[2];
            const int c2 = (wx * g.nc[1] + wy) * g.nc[2] + wz;
            const int cnt = d.ccount[c2];
            for (int s2 = 0; s2 < cnt; s2++) {
              const size_t sj = (size_t)c2 * g.cap + s2;
              const int j = d.cbead[sj];
              if (j < i) continue;
              const double4 pj = d.cpos[sj];
              const double dx = pj.x + qx * g.L[0] - pw[0];
              const double dy = pj.y + qy * g.L[1] - pw[1];
              const double dz = pj.z + qz * g.L[2] - pw[2];
              const double r2 = dx * dx + dy * dy + dz * dz;
              if (r2 > g.rs2) continue;
              const double r = sqrt(r2);
              if (j > i && g.use_pair && (g.npbc == 3 || qz == 0)) elj += lj_energy(r, ljrow[d.ctype[sj]]);
              if (g.use_ewald) {
                const double qq = p.w * pj.w;
                if (qq != 0 && r > 0 && r <= g.rc) {
                  const double e = g.lB * qq * erfc(g.sqrt_alpha * r) / r;
                  ere += (j == i) ? 0.5 * e : e;
                }
              }
            }
          }
        }
      }
    }
  }
  const double a = block_sum<kBlock>(elj, sh);
  const double c = block_sum<kBlock>(ere, sh);
  if (threadIdx.x == 0) {
    d.fpart[2 * blockIdx.x] = a;
    d.fpart[2 * blockIdx.x + 1] = c;
  }
}

// S(k) from scratch, bead chunk `blockIdx.y`: direct FP64 sincos per (k, bead)
__global__ __launch_bounds__(kBlock) void k_full_sk_part(Geo g, Dev d, int chunk) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= g.n_kh) return;
  const int4 l = d.kl[k];
  const double kx = (double)(l.x * 2) * kPi / g.L[0];
  const double ky = (double)(l.y * 2) * kPi / g.L[1];
  const double kz = (double)(l.z * 2) * kPi / g.L[2];
  const int j0 = blockIdx.y * chunk, j1 = min(g.n, j0 + chunk);
  double sr = 0, si = 0;
  for (int j = j0; j < j1; j++) {
    const double qj = d.q[j];
    if (qj == 0) continue;
    const double xw = wrap_coord(d.x[j], g.L[0], g.inv_L[0]);
    const double yw = wrap_coord(d.y[j], g.L[1], g.inv_L[1]);
    const double zw = wrap_coord(d.z[j], g.L[2], g.inv_L[2]);
    double s, c;
    sincos(kx * xw + ky * yw + kz * zw, &s, &c);
    sr += qj * c;
    si += qj * s;
  }
  d.sk_part[(size_t)blockIdx.y * g.n_kh + k] = make_double2(sr, si);
}

__global__ __launch_bounds__(kBlock) void k_full_sk_reduce(Geo g, Dev d, int nchunk,
                                                           double* epart) {
  __shared__ double sh[kBlock / kWave];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  double e = 0;
  if (k < g.n_kh) {
    double2 s = make_double2(0, 0);
    for (int c = 0; c < nchunk; c++) {
      const double2 v = d.sk_part[(size_t)c * g.n_kh + k];
      s.x += v.x;
      s.y += v.y;
    }
    d.Sk[k] = s;
    e = d.ck[k] * (s.x * s.x + s.y * s.y);
  }
  const double r = block_sum<kBlock>(e, sh);
  if (threadIdx.x == 0) epart[blockIdx.x] = r;
}

// Per-molecule springs, wall, self energy and the dipole moment; then the
// final reduction of all from-scratch partials into the control block.
__global__ __launch_bounds__(kBlock) void k_full_small(Geo g, Dev d, int nrealb, int nkb,
                                                       const double* epart_k, int set_running) {
  __shared__ double sh[kBlock / kWave];
  __shared__ double sbuf[kBlock];
  Ctl* C = d.ctl;
  double eb_all = 0;
  if (g.use_bond) {
    for (int m = 0; m < g.n_mols; m++) {
      double e = 0;
      if (d.mol_start[m + 1] - d.mol_start[m] > 1) e = molecule_spring(g, d, m, false, sbuf);
      if (threadIdx.x == 0) {
        d.bond_cur[m] = e;
        eb_all += e;
      }
    }
  }
  double ew = 0, es = 0, mz = 0;
  for (int i = threadIdx.x; i < g.n; i += kBlock) {
    const double qi = d.q[i];
    if (g.use_ext) ew += wall_energy(d.z[i], d.wall[d.type[i]], g);
    es += g.self_coef * qi * qi;
    mz += qi * d.z[i];
  }
  double vl = 0, vr = 0, vk = 0;
  for (int b = threadIdx.x; b < nrealb; b += kBlock) { vl += d.fpart[2 * b]; vr += d.fpart[2 * b + 1]; }
  for (int b = threadIdx.x; b < nkb; b += kBlock) vk += epart_k[b];
  const double ewt = block_sum<kBlock>(ew, sh);
  const double est = block_sum<kBlock>(es, sh);
  const double mzt = block_sum<kBlock>(mz, sh);
  const double lt = block_sum<kBlock>(vl, sh);
  const double rt = block_sum<kBlock>(vr, sh);
  const double kt = block_sum<kBlock>(vk, sh);
  if (threadIdx.x == 0) {
    double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g.use_pair) f[T_LJ] = lt;
    if (g.use_ewald) {
      f[T_REAL] = rt;
      f[T_KSPACE] = kt;
      f[T_SELF] = est;
      if (g.dipole) f[T_DIPOLE] = g.dip_coef * mzt * mzt;
    }
    if (g.use_bond) f[T_BOND] = eb_all;
    if (g.use_ext) f[T_WALL] = ewt;
    for (int k = 0; k < 8; k++) C->full[k] = f[k];
    if (set_running) {
      C->Mz = mzt;
      for (int k = 0; k < 8; k++) C->running[k] = f[k];
      C->cur_dipl = C->trial_dipl = f[T_DIPOLE];
    }
  }
}

}  // namespace

// ============================================================ host side
struct plum_ctx {
  plum_params p{};
  Geo g{};
  plum_info info{};
  Dev d{};
  hipStream_t st = nullptr;
  int maxm = 1;          // capacity of the moved list (largest molecule)
  int nkb = 0;           // blocks of the k-space kernels
  int nmmb = 1;          // blocks of the moved-moved kernel
  int nchunk = 1;        // bead chunks of the S(k) build
  int chunk = 512;
  int nrealb = 1;        // blocks of the full real-space kernel
  double* epart_k = nullptr;
  std::vector<int> mol_start;
  std::vector<int> type_h;
  std::vector<double> q_h;
  // pinned staging
  char* h_in = nullptr;
  Ctl* h_ctl = nullptr;
  size_t in_bytes = 0;
  char* d_in = nullptr;
  bool pending = false;
  bool profiling = false;
  std::vector<void*> dev_allocs;
};

namespace {

template <typename T>
int dalloc(plum_ctx* c, T** ptr, size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * count);
  if (e != hipSuccess) return fail(PLUM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->dev_allocs.push_back(p);
  *ptr = static_cast<T*>(p);
  return PLUM_OK;
}

#define CHK(x)                 \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != PLUM_OK) return rc_; \
  } while (0)

// potential_ewald_coul.cc:43-75: slab padding and the derived cutoffs.
void derive_cutoffs(const plum_params& p, plum_info& info) {
  double box[3] = {p.box[0], p.box[1], p.box[2]};
  if (p.dipole_correction) {
    const double min_padding = 150;
    if (box[2] * kDiCorrection > min_padding) box[2] = box[2] * kDiCorrection;
    else box[2] += min_padding;
  }
  for (int i = 0; i < 3; i++) info.box_e[i] = box[i];
  info.vol_e = box[0] * box[1] * box[2];
  if (!p.use_ewald) return;
  const double lB = p.lB, alpha = p.alpha;
  double real_cutoff = 1;
  while (0.5 * lB * 1 * 1 * erfc(sqrt(alpha) * real_cutoff) / real_cutoff > kEwaldCutoff)
    real_cutoff += 1;
  info.real_cutoff = real_cutoff;
  for (int i = 0; i < 3; i++) info.real_cell[i] = (int)ceil(real_cutoff / box[i]);
  double repl_cutoff = alpha;
  const double min_box_vol = pow(box[0], 3);
  while (lB * 1 * 1 / (2 * kPi * min_box_vol) * (4 * kPi * kPi) * exp(-repl_cutoff / (4 * alpha)) /
             repl_cutoff >
         kEwaldCutoff)
    repl_cutoff += alpha;
  info.repl_cutoff = repl_cutoff;
  for (int i = 0; i < 3; i++) info.repl_cell[i] = (int)ceil(sqrt(repl_cutoff) * box[i] / (2 * kPi));
}

double host_wrap(double x, double L, double invL) {
  double w = x - L * floor(x * invL);
  if (w >= L) w -= L;
  if (w < 0) w += L;
  return w;
}

int host_cell(double xw, double inv_h, int n) {
  int c = (int)floor(xw * inv_h);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// Builds the cell list from the current device positions (host counting
// sort in bead order, so slot order is deterministic) and uploads it.
int build_cells(plum_ctx* c, int min_cap) {
  const Geo& g = c->g;
  const int n = g.n;
  std::vector<double> x(n), y(n), z(n);
  HIPCHK(hipMemcpyAsync(x.data(), c->d.x, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(y.data(), c->d.y, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(z.data(), c->d.z, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  std::vector<int> cell(n), count(g.ncell, 0);
  std::vector<double> w(3 * (size_t)n);
  for (int i = 0; i < n; i++) {
    const double p[3] = {x[i], y[i], z[i]};
    int cc = 0;
    for (int k = 0; k < 3; k++) {
      w[3 * (size_t)i + k] = host_wrap(p[k], g.L[k], g.inv_L[k]);
      cc = cc * g.nc[k] + host_cell(w[3 * (size_t)i + k], g.inv_h[k], g.nc[k]);
    }
    cell[i] = cc;
    count[cc]++;
  }
  int maxc = *std::max_element(count.begin(), count.end());
  int cap = std::max(std::max(16, min_cap), ((maxc * 2 + 8 + 7) / 8) * 8);
  const size_t slots = (size_t)g.ncell * cap;
  if (cap != c->g.cap || c->d.cbead == nullptr) {
    // (re)allocate slot arrays; old ones stay in dev_allocs until destroy
    CHK(dalloc(c, &c->d.cbead, slots));
    CHK(dalloc(c, &c->d.cpos, slots));
    CHK(dalloc(c, &c->d.ctype, slots));
    c->g.cap = cap;
  }
  std::vector<int> cbead(slots, -1), ctype(slots, 0), bslot(n);
  std::vector<double4> cpos(slots, make_double4(0, 0, 0, 0));
  std::fill(count.begin(), count.end(), 0);
  for (int i = 0; i < n; i++) {
    const int cc = cell[i];
    const int s = count[cc]++;
    const size_t si = (size_t)cc * cap + s;
    cbead[si] = i;
    ctype[si] = c->type_h[i];
    cpos[si] = make_double4(w[3 * (size_t)i], w[3 * (size_t)i + 1], w[3 * (size_t)i + 2], c->q_h[i]);
    bslot[i] = s;
  }
  HIPCHK(hipMemcpyAsync(c->d.ccount, count.data(), sizeof(int) * g.ncell, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.cbead, cbead.data(), sizeof(int) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.ctype, ctype.data(), sizeof(int) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.cpos, cpos.data(), sizeof(double4) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.bcell, cell.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.bslot, bslot.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->info.cell_cap = cap;
  c->nrealb = (int)((slots + kBlock - 1) / kBlock);
  // full-energy partials sized for the current slot count
  CHK(dalloc(c, &c->d.fpart, 2 * (size_t)c->nrealb));
  return PLUM_OK;
}

// From-scratch energy on the device; with set_running the running totals,
// dipole state and per-molecule spring cache are reset to it.
int run_full(plum_ctx* c, int set_running, double out[8]) {
  const Geo& g = c->g;
  hipLaunchKernelGGL(k_full_real, dim3(c->nrealb), dim3(kBlock), 0, c->st, g, c->d);
  if (g.use_ewald && g.n_kh > 0) {
    hipLaunchKernelGGL(k_full_sk_part, dim3(c->nkb, c->nchunk), dim3(kBlock), 0, c->st, g, c->d,
                       c->chunk);
    hipLaunchKernelGGL(k_full_sk_reduce, dim3(c->nkb), dim3(kBlock), 0, c->st, g, c->d, c->nchunk,
                       c->epart_k);
  }
  hipLaunchKernelGGL(k_full_small, dim3(1), dim3(kBlock), 0, c->st, g, c->d, c->nrealb,
                     (g.use_ewald ? c->nkb : 0), c->epart_k, set_running);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int k = 0; k < 8; k++) out[k] = c->h_ctl->full[k];
  return PLUM_OK;
}

void to_terms(const double t[8], plum_terms* o) {
  o->lj = t[T_LJ]; o->real = t[T_REAL]; o->kspace = t[T_KSPACE]; o->self = t[T_SELF];
  o->dipole = t[T_DIPOLE]; o->bond = t[T_BOND]; o->wall = t[T_WALL];
}

int validate(const plum_params* p, int n, const double* xyz, const double* q, const int* type,
             int n_mols, const int* ms) {
  if (!p || !xyz || !q || !type || !ms) return fail(PLUM_E_ARG, "null argument");
  if (n <= 0 || n_mols <= 0) return fail(PLUM_E_ARG, "empty system");
  if (p->npbc != 2 && p->npbc != 3)
    return fail(PLUM_E_ARG, "npbc must be 2 or 3 (simulation.cc:152-156)");
  if (p->npbc == 3 && p->dipole_correction)
    return fail(PLUM_E_UNSUPPORTED, "dipole correction with 3 periodic dimensions");
  if (!(p->beta > 0)) return fail(PLUM_E_ARG, "beta must be > 0 (simulation.cc:142-146)");
  for (int k = 0; k < 3; k++)
    if (!(p->box[k] > 0) || !std::isfinite(p->box[k])) return fail(PLUM_E_ARG, "box lengths must be > 0");
  if (p->n_types <= 0 || p->n_types > PLUM_MAX_TYPES) return fail(PLUM_E_ARG, "n_types out of range");
  if (p->use_ewald && !(p->alpha > 0 && p->lB >= 0)) return fail(PLUM_E_ARG, "Ewald alpha must be > 0");
  if (ms[0] != 0 || ms[n_mols] != n) return fail(PLUM_E_ARG, "mol_start must span 0..n_beads");
  for (int m = 0; m < n_mols; m++)
    if (ms[m + 1] <= ms[m]) return fail(PLUM_E_ARG, "empty or unordered molecule");
  for (int i = 0; i < n; i++) {
    if (type[i] < 0 || type[i] >= p->n_types) return fail(PLUM_E_ARG, "bead type out of range");
    for (int k = 0; k < 3; k++)
      if (!std::isfinite(xyz[3 * (size_t)i + k])) return fail(PLUM_E_ARG, "non-finite coordinate");
    if (!std::isfinite(q[i])) return fail(PLUM_E_ARG, "non-finite charge");
  }
  return PLUM_OK;
}

}  // namespace

extern "C" {

const char* plum_last_error(void) { return g_err.c_str(); }

int plum_create(const plum_params* params, int n, const double* xyz, const double* q,
                const int* type, int n_mols, const int* ms, plum_ctx** out) {
  if (!out) return fail(PLUM_E_ARG, "null output");
  *out = nullptr;
  CHK(validate(params, n, xyz, q, type, n_mols, ms));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PLUM_E_HIP, "no HIP device available");
  plum_ctx* c = new plum_ctx();
  c->p = *params;
  const plum_params& p = c->p;
  int rc = PLUM_OK;
  auto bail = [&](int code) {
    plum_destroy(c);
    return code;
  };
  plum_info& info = c->info;
  derive_cutoffs(p, info);
  Geo& g = c->g;
  for (int k = 0; k < 3; k++) {
    g.L[k] = info.box_e[k];
    g.Lsim[k] = p.box[k];
    g.inv_L[k] = 1.0 / g.L[k];
  }
  g.npbc = p.npbc;
  g.n = n;
  g.n_mols = n_mols;
  g.beta = p.beta;
  g.use_pair = p.use_pair;
  g.use_ewald = p.use_ewald;
  g.use_bond = p.use_bond;
  g.use_ext = p.use_ext;
  g.n_types = p.n_types;
  g.rc = p.use_ewald ? info.real_cutoff : 0;
  g.sqrt_alpha = p.use_ewald ? sqrt(p.alpha) : 0;
  g.lB = p.lB;
  g.self_coef = p.use_ewald ? -p.lB * sqrt(p.alpha / kPi) : 0;
  g.dip_coef = p.lB * 2 * kPi / (info.vol_e / 1.0);
  for (int k = 0; k < 3; k++) g.real_cell[k] = info.real_cell[k];
  g.k_bond = p.k_bond;
  g.r0 = p.r0;
  g.wall_cut = p.wall_cut;
  g.wall_r3_k213 = pow((1.0 / k213), 3);
  g.wall_r3_cut = pow((1.0 / p.wall_cut), 3);
  g.dipole = p.use_ewald && p.dipole_correction;
  g.dipole_exact = p.dipole_exact;

  // LJ pair table (potential_truncated_lj.cc:49-85, Lorentz-Berthelot)
  std::vector<LJPair> lj((size_t)p.n_types * p.n_types);
  double lj_max = 0;
  for (int a = 0; a < p.n_types; a++)
    for (int b = 0; b < p.n_types; b++) {
      LJPair& e = lj[(size_t)a * p.n_types + b];
      e.sigma = (p.lj_sigma[a] + p.lj_sigma[b]) / 2;
      e.eps = sqrt((p.lj_eps[a] * p.lj_eps[b]));
      if (p.lj_cutoff < 0) {
        const double r6_ref = pow((1.0 / k216), 6);
        e.shift = 4 * e.eps * (r6_ref * r6_ref - r6_ref);
        e.cut = k216 * e.sigma;
      } else {
        const double r6_ref = pow((e.sigma / p.lj_cutoff), 6);
        e.shift = 4 * e.eps * (r6_ref * r6_ref - r6_ref);
        e.cut = p.lj_cutoff;
      }
      if (p.use_pair) lj_max = std::max(lj_max, e.cut);
    }
  for (int k = 0; k < p.npbc; k++)
    if (p.use_pair && !(p.box[k] > 2 * lj_max)) {
      delete c;
      return fail(PLUM_E_UNSUPPORTED, "periodic box shorter than twice the LJ cutoff");
    }
  std::vector<WallType> wall(p.n_types);
  for (int t = 0; t < p.n_types; t++) {
    wall[t].sigma = p.wall_sigma[t];
    wall[t].eps = p.wall_eps[t];
    wall[t].kind = p.type_kind[t];
    wall[t].pad = 0;
  }
  g.rs = std::max(g.rc, lj_max);
  if (g.rs <= 0) g.rs = 1.0;
  g.rs2 = g.rs * g.rs;

  // cell grid: cell edge ~ rs / div
  const int div = p.cell_div > 0 ? p.cell_div : 2;
  for (int k = 0; k < 3; k++) {
    int nc = (int)floor(g.L[k] / (g.rs / div));
    nc = std::max(1, std::min(nc, 256));
    g.nc[k] = nc;
    g.h[k] = g.L[k] / nc;
    g.inv_h[k] = 1.0 / g.h[k];
    info.n_cells[k] = nc;
  }
  g.ncell = g.nc[0] * g.nc[1] * g.nc[2];

  // half-space k list (potential_ewald_coul.cc:78-115; k and -k paired)
  std::vector<int4> kl;
  std::vector<double> ck;
  info.n_k_full = 0;
  if (p.use_ewald) {
    const int* n3 = info.repl_cell;
    g.nkx = n3[0]; g.nky = n3[1]; g.nkz = n3[2];
    const double pref0 = p.lB / (kPi * info.vol_e) * (4 * kPi * kPi);
    for (int lx = -n3[0]; lx <= n3[0]; lx++)
      for (int ly = -n3[1]; ly <= n3[1]; ly++)
        for (int lz = -n3[2]; lz <= n3[2]; lz++) {
          const double kx = (double)(lx * 2) * kPi / g.L[0];
          const double ky = (double)(ly * 2) * kPi / g.L[1];
          const double kz = (double)(lz * 2) * kPi / g.L[2];
          const double k2 = kx * kx + ky * ky + kz * kz;
          if (!(k2 > 0 && k2 <= info.repl_cutoff)) continue;
          info.n_k_full++;
          const bool half = lx > 0 || (lx == 0 && (ly > 0 || (ly == 0 && lz > 0)));
          if (!half) continue;
          kl.push_back(make_int4(lx, ly, lz, 0));
          ck.push_back(pref0 * (exp(-k2 / (4 * p.alpha)) / k2));
        }
    g.tab_w = (g.nkx + 1) + (2 * g.nky + 1) + (2 * g.nkz + 1);
  }
  g.n_kh = (int)kl.size();
  info.n_k_half = g.n_kh;
  c->nkb = std::max(1, (g.n_kh + kBlock - 1) / kBlock);

  c->mol_start.assign(ms, ms + n_mols + 1);
  c->type_h.assign(type, type + n);
  c->q_h.assign(q, q + n);
  int maxm = 1;
  for (int m = 0; m < n_mols; m++) maxm = std::max(maxm, ms[m + 1] - ms[m]);
  c->maxm = maxm;
  info.n_beads = n;
  info.n_mols = n_mols;
  info.max_mol_len = maxm;
  c->nmmb = std::max(1, std::min(64, (maxm * (maxm - 1) / 2 + kBlock - 1) / kBlock));
  c->nchunk = (n + c->chunk - 1) / c->chunk;

  if (hipStreamCreate(&c->st) != hipSuccess) {
    delete c;
    return fail(PLUM_E_HIP, "hipStreamCreate failed");
  }
  Dev& d = c->d;
  std::vector<double> hx(n), hy(n), hz(n);
  std::vector<int> mol_of(n);
  for (int i = 0; i < n; i++) {
    hx[i] = xyz[3 * (size_t)i]; hy[i] = xyz[3 * (size_t)i + 1]; hz[i] = xyz[3 * (size_t)i + 2];
  }
  for (int m = 0; m < n_mols; m++)
    for (int i = ms[m]; i < ms[m + 1]; i++) mol_of[i] = m;
#define ALLOC(ptr, cnt) if ((rc = dalloc(c, &ptr, cnt)) != PLUM_OK) return bail(rc)
  ALLOC(d.x, n); ALLOC(d.y, n); ALLOC(d.z, n); ALLOC(d.q, n);
  ALLOC(d.type, n); ALLOC(d.mol_of, n); ALLOC(d.mol_start, n_mols + 1);
  ALLOC(d.bcell, n); ALLOC(d.bslot, n); ALLOC(d.moved, n); ALLOC(d.mv_slot, n);
  ALLOC(d.ccount, g.ncell);
  ALLOC(d.kl, g.n_kh); ALLOC(d.ck, g.n_kh); ALLOC(d.Sk, g.n_kh); ALLOC(d.dS, g.n_kh);
  ALLOC(d.tab, (size_t)maxm * 2 * std::max(1, g.tab_w));
  ALLOC(d.lj, lj.size()); ALLOC(d.wall, wall.size());
  ALLOC(d.bond_cur, n_mols);
  ALLOC(d.mv_old, maxm); ALLOC(d.mv_new, maxm); ALLOC(d.mv_type, maxm);
  ALLOC(d.ctl, 1);
  ALLOC(d.part_k, c->nkb); ALLOC(d.part_r, 2 * maxm); ALLOC(d.part_mm, c->nmmb);
  ALLOC(d.sk_part, (size_t)c->nchunk * std::max(1, g.n_kh));
  ALLOC(c->epart_k, c->nkb);
  // move record: header + idx[maxm] + xyz[3 maxm], one contiguous upload
  c->in_bytes = 16 + sizeof(int) * maxm + ((sizeof(int) * maxm) % 8) + sizeof(double) * 3 * maxm;
  {
    char* din = nullptr;
    ALLOC(din, c->in_bytes);
    c->d_in = din;
    d.rec = reinterpret_cast<MoveRec*>(din);
    d.mv_idx = reinterpret_cast<int*>(din + 16);
    d.mv_xyz = reinterpret_cast<double*>(din + 16 + sizeof(int) * maxm + ((sizeof(int) * maxm) % 8));
  }
#undef ALLOC
  if (hipHostMalloc((void**)&c->h_in, c->in_bytes, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess)
    return bail(fail(PLUM_E_NOMEM, "hipHostMalloc failed"));
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return PLUM_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->st));
    return PLUM_OK;
  };
  if ((rc = up(d.x, hx.data(), sizeof(double) * n)) || (rc = up(d.y, hy.data(), sizeof(double) * n)) ||
      (rc = up(d.z, hz.data(), sizeof(double) * n)) || (rc = up(d.q, q, sizeof(double) * n)) ||
      (rc = up(d.type, type, sizeof(int) * n)) || (rc = up(d.mol_of, mol_of.data(), sizeof(int) * n)) ||
      (rc = up(d.mol_start, ms, sizeof(int) * (n_mols + 1))) ||
      (rc = up(d.kl, kl.data(), sizeof(int4) * kl.size())) ||
      (rc = up(d.ck, ck.data(), sizeof(double) * ck.size())) ||
      (rc = up(d.lj, lj.data(), sizeof(LJPair) * lj.size())) ||
      (rc = up(d.wall, wall.data(), sizeof(WallType) * wall.size())))
    return bail(rc);
  if (hipMemsetAsync(d.moved, 0, n, c->st) != hipSuccess ||
      hipMemsetAsync(d.mv_slot, 0xff, sizeof(int) * n, c->st) != hipSuccess ||
      hipMemsetAsync(d.ctl, 0, sizeof(Ctl), c->st) != hipSuccess ||
      hipMemsetAsync(c->d_in, 0, c->in_bytes, c->st) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "hipMemsetAsync failed"));
  if ((rc = build_cells(c, 0)) != PLUM_OK) return bail(rc);
  double f[8];
  if ((rc = run_full(c, 1, f)) != PLUM_OK) return bail(rc);
  *out = c;
  return PLUM_OK;
}

int plum_destroy(plum_ctx* c) {
  if (!c) return PLUM_OK;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (void* p : c->dev_allocs) (void)hipFree(p);
  if (c->h_in) (void)hipHostFree(c->h_in);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return PLUM_OK;
}

int plum_get_info(const plum_ctx* c, plum_info* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  *out = c->info;
  return PLUM_OK;
}

int plum_energy_full(plum_ctx* c, plum_terms* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  double f[8];
  CHK(run_full(c, 0, f));
  to_terms(f, out);
  return PLUM_OK;
}

int plum_energy_difference(plum_ctx* c, int mol, int n_moved, const int* idx,
                           const double* trial_xyz, plum_terms* dE, double* total) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (c->pending) return fail(PLUM_E_STATE, "previous trial move not finalized");
  if (mol < 0 || mol >= c->g.n_mols) return fail(PLUM_E_ARG, "molecule index out of range");
  const int a = c->mol_start[mol], b = c->mol_start[mol + 1];
  if (n_moved < 0 || n_moved > b - a) return fail(PLUM_E_ARG, "n_moved out of range");
  if (n_moved > 0 && (!idx || !trial_xyz)) return fail(PLUM_E_ARG, "null moved list");
  std::vector<char> seen(b - a, 0);
  for (int k = 0; k < n_moved; k++) {
    if (idx[k] < a || idx[k] >= b) return fail(PLUM_E_ARG, "moved bead not in the molecule");
    if (seen[idx[k] - a]++) return fail(PLUM_E_ARG, "moved bead listed twice");
    for (int e = 0; e < 3; e++)
      if (!std::isfinite(trial_xyz[3 * k + e])) return fail(PLUM_E_ARG, "non-finite trial coordinate");
  }
  MoveRec* r = reinterpret_cast<MoveRec*>(c->h_in);
  r->n_moved = n_moved;
  r->mol = mol;
  r->move_type = -1;
  r->valid = 1;
  const size_t off_xyz = 16 + sizeof(int) * c->maxm + ((sizeof(int) * c->maxm) % 8);
  std::memcpy(c->h_in + 16, idx, sizeof(int) * n_moved);
  std::memcpy(c->h_in + off_xyz, trial_xyz, sizeof(double) * 3 * n_moved);
  HIPCHK(hipMemcpyAsync(c->d_in, c->h_in, 16 + sizeof(int) * c->maxm, hipMemcpyHostToDevice, c->st));
  if (n_moved > 0)
    HIPCHK(hipMemcpyAsync(c->d_in + off_xyz, c->h_in + off_xyz, sizeof(double) * 3 * n_moved,
                          hipMemcpyHostToDevice, c->st));
  const Geo& g = c->g;
  const int prep_work = std::max(n_moved, n_moved * 2 * std::max(1, g.tab_w));
  const int prep_blocks = std::max(1, std::min(1024, (prep_work + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_prep, dim3(prep_blocks), dim3(kBlock), 0, c->st, g, c->d);
  if (g.use_ewald && g.n_kh > 0)
    hipLaunchKernelGGL(k_kspace_delta, dim3(c->nkb), dim3(kBlock), 0, c->st, g, c->d);
  if (n_moved > 0)
    hipLaunchKernelGGL(k_real_delta, dim3(2 * n_moved), dim3(kBlock), 0, c->st, g, c->d);
  const int npair = n_moved * (n_moved - 1) / 2;
  const int nmmb = std::max(1, std::min(c->nmmb, (npair + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_mm_delta, dim3(nmmb), dim3(kBlock), 0, c->st, g, c->d);
  hipLaunchKernelGGL(k_epilogue, dim3(1), dim3(kBlock), 0, c->st, g, c->d,
                     (g.use_ewald && g.n_kh > 0) ? c->nkb : 0, 2 * n_moved, nmmb, 0, 0ULL);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->pending = true;
  if (dE) to_terms(c->h_ctl->terms, dE);
  if (total) *total = c->h_ctl->total;
  return PLUM_OK;
}

int plum_finalize(plum_ctx* c, int accept) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (!c->pending) return fail(PLUM_E_STATE, "no pending trial move");
  const Geo& g = c->g;
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(kBlock), 0, c->st, g, c->d, accept ? 1 : 0, 0);
  if (g.use_ewald && g.n_kh > 0 && accept)
    hipLaunchKernelGGL(k_commit_sk, dim3(c->nkb), dim3(kBlock), 0, c->st, g, c->d, 1, 0);
  HIPCHK(hipGetLastError());
  c->pending = false;
  if (accept) {
    HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->h_ctl->overflow) {
      // a cell ran out of slots: rebuild with twice the capacity
      int zero = 0;
      HIPCHK(hipMemcpyAsync(&c->d.ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
      CHK(build_cells(c, 2 * c->g.cap));
    }
  }
  return PLUM_OK;
}

int plum_totals(plum_ctx* c, plum_terms* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  to_terms(c->h_ctl->running, out);
  return PLUM_OK;
}

int plum_resync(plum_ctx* c, plum_terms* drift) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  double before[8];
  for (int k = 0; k < 8; k++) before[k] = c->h_ctl->running[k];
  int zero = 0;
  HIPCHK(hipMemcpyAsync(&c->d.ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
  CHK(build_cells(c, c->g.cap));
  double f[8];
  CHK(run_full(c, 1, f));
  if (drift) {
    double dd[8];
    for (int k = 0; k < 8; k++) dd[k] = before[k] - f[k];
    to_terms(dd, drift);
  }
  return PLUM_OK;
}

int plum_get_positions(plum_ctx* c, double* xyz) {
  if (!c || !xyz) return fail(PLUM_E_ARG, "null argument");
  const int n = c->g.n;
  std::vector<double> x(n), y(n), z(n);
  HIPCHK(hipMemcpyAsync(x.data(), c->d.x, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(y.data(), c->d.y, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(z.data(), c->d.z, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int i = 0; i < n; i++) {
    xyz[3 * (size_t)i] = x[i]; xyz[3 * (size_t)i + 1] = y[i]; xyz[3 * (size_t)i + 2] = z[i];
  }
  return PLUM_OK;
}

int plum_set_profiling(plum_ctx* c, int on) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  c->profiling = on != 0;
  return PLUM_OK;
}

int plum_mc_run(plum_ctx* c, const plum_mc_params* mc, long long n_moves, plum_mc_stats* stats) {
  (void)mc; (void)n_moves; (void)stats;
  if (!c) return fail(PLUM_E_ARG, "null context");
  return fail(PLUM_E_UNSUPPORTED, "device-resident MC not built yet");
}

int plum_kernel_times(plum_ctx* c, int cap, const char** names, double* avg_us, long long* count) {
  (void)cap; (void)names; (void)avg_us; (void)count;
  if (!c) return fail(PLUM_E_ARG, "null context");
  return 0;
}

}  // extern "C"
