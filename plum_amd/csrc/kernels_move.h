// Kernels of one trial move: proposal (device loop), phase tables, k-space
// dS(k)/dE, real space + LJ over the cell list, epilogue (bonded, wall,
// dipole, early exits, Metropolis) and the commit.
#pragma once
#include "engine_state.h"

namespace plum {

constexpr int kKGroups = 16;            // bead groups per k-space block
constexpr int kKPerBlock = kBlock / kKGroups;
constexpr int kMaxChainDev = 2048;      // longest molecule the device loop moves

// ------------------------------------------------------------ move list
// Host path: the caller's moved list -> old/new positions, masks.
__global__ void k_host_moves(Geo g, Dev d) {
  const int nm = d.rec->n_moved;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
    const int i = d.mv_idx[k];
    d.moved[i] = 1;
    d.mv_slot[i] = k;
    const double qi = d.q[i];
    d.mv_old[k] = make_double4(d.x[i], d.y[i], d.z[i], qi);
    d.mv_new[k] = make_double4(d.mv_xyz[3 * k], d.mv_xyz[3 * k + 1], d.mv_xyz[3 * k + 2], qi);
    d.mv_type[k] = d.type[i];
  }
}

// Phase tables e^{i k_l x} for l = 0..n_d of each dimension d of every moved
// bead at its old and new position, with k_l = (l*2) kPi / L exactly as
// potential_ewald_coul.cc:89-96 forms it (FP64 sincos).  Negative l is read
// as the conjugate (tab_phase).
__global__ void k_tables(Geo g, Dev d) {
  const int nm = d.rec->n_moved;
  const int total = nm * 2 * g.tab_w;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const int row = e / g.tab_w, off = e - row * g.tab_w;
    const int m = row >> 1, on = row & 1;
    int dim, l;
    if (off <= g.nkx) { dim = 0; l = off; }
    else if (off <= g.nkx + 1 + g.nky) { dim = 1; l = off - (g.nkx + 1); }
    else { dim = 2; l = off - (g.nkx + 1) - (g.nky + 1); }
    const double4 p = on ? d.mv_new[m] : d.mv_old[m];
    const double xv = dim == 0 ? p.x : (dim == 1 ? p.y : p.z);
    const double xw = wrap_coord(xv, g.L[dim], g.inv_L[dim]);
    const double kd = (double)(l * 2) * kPi / g.L[dim];
    double s, c;
    sincos(kd * xw, &s, &c);
    d.tab[(size_t)row * g.tab_w + off] = make_double2(c, s);
  }
}

// Table offsets of k vector l (x, y, z segments of l >= 0) and whether its y
// and z phases are conjugated (negative l).
struct TabIdx {
  int ox, oy, oz;
  bool ny, nz;
};

__device__ __forceinline__ TabIdx tab_index(const Geo& g, int4 l) {
  TabIdx t;
  t.ox = l.x;
  t.oy = g.nkx + 1 + abs(l.y);
  t.oz = g.nkx + 1 + g.nky + 1 + abs(l.z);
  t.ny = l.y < 0;
  t.nz = l.z < 0;
  return t;
}

// e^{i k.r} of one table row: e^{-i l theta} = conj(e^{i l theta}).
__device__ __forceinline__ double2 tab_phase(const double2* t, const TabIdx& i) {
  double2 y = t[i.oy], z = t[i.oz];
  if (i.ny) y.y = -y.y;
  if (i.nz) z.y = -z.y;
  return cmul(cmul(t[i.ox], y), z);
}

// ------------------------------------------------------------ k space
// dS(k) = sum_m q_m (e^{ik.r'_m} - e^{ik.r_m}); dE_k = c_k (|S+dS|^2 - |S|^2).
// Block = 16 k vectors x 16 groups of moved beads; the groups are reduced in
// LDS.  A still-pending S(k) += dS(k) of the previous accepted move is applied
// first by the thread that owns k (device loop).  FP64 in both precision
// modes: it is a small part of the per-move work and keeps S(k) exact.
struct KSmem {
  double2 s_ds[kKGroups][kKPerBlock];
  double sh[kBlock / kWave];
};

__device__ void kspace_block(const Geo& g, const Dev& d, int blk, KSmem& sm) {
  double2 (&s_ds)[kKGroups][kKPerBlock] = sm.s_ds;
  double* sh = sm.sh;
  const int kk = threadIdx.x % kKPerBlock, grp = threadIdx.x / kKPerBlock;
  const int k = blk * kKPerBlock + kk;
  const int nm = d.rec->n_moved;
  double2 dS = make_double2(0, 0);
  if (k < g.n_kh) {
    const TabIdx ti = tab_index(g, d.kl[k]);
#pragma unroll 2
    for (int m = grp; m < nm; m += kKGroups) {
      const double qm = d.mv_old[m].w;
      const double2* to = d.tab + (size_t)(2 * m) * g.tab_w;
      const double2* tn = to + g.tab_w;
      const double2 eo = tab_phase(to, ti);
      const double2 en = tab_phase(tn, ti);
      dS.x += qm * (en.x - eo.x);
      dS.y += qm * (en.y - eo.y);
    }
  }
  s_ds[grp][kk] = dS;
  __syncthreads();
  double e = 0;
  if (grp == 0 && k < g.n_kh) {
    double2 t = s_ds[0][kk];
#pragma unroll
    for (int r = 1; r < kKGroups; r++) { t.x += s_ds[r][kk].x; t.y += s_ds[r][kk].y; }
    double2 S = d.Sk[k];
    if (d.ctl->sk_pending) {
      const double2 prev = d.dS[k];
      S.x += prev.x;
      S.y += prev.y;
      d.Sk[k] = S;
    }
    e = d.ck[k] * (2.0 * (S.x * t.x + S.y * t.y) + t.x * t.x + t.y * t.y);
    d.dS[k] = t;
  }
  const double r = block_sum<kBlock>(e, sh);
  if (threadIdx.x == 0) d.part_k[blk] = r;
}

// Applies a pending S(k) += dS(k) (end of a device run).
__global__ void k_flush_sk(Geo g, Dev d) {
  if (!d.ctl->sk_pending) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < g.n_kh) {
    double2 s = d.Sk[k];
    const double2 ds = d.dS[k];
    s.x += ds.x;
    s.y += ds.y;
    d.Sk[k] = s;
  }
}

// Clears the pending S(k) flag after k_flush_sk (single thread).
__global__ void k_clear_sk_pending(Dev d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d.ctl->sk_pending = 0;
}

// Host path: S(k) += dS(k) right away when the caller accepts.
__global__ void k_commit_sk(Geo g, Dev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < g.n_kh) {
    double2 s = d.Sk[k];
    const double2 ds = d.dS[k];
    s.x += ds.x;
    s.y += ds.y;
    d.Sk[k] = s;
  }
}

// ------------------------------------------------------------ real space
// Inclusive scan of n <= 1024 ints in LDS (Hillis-Steele, blockDim = 256).
__device__ void lds_scan(int* a, int n) {
  for (int off = 1; off < n; off <<= 1) {
    int v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = threadIdx.x + r * kBlock;
      v[r] = (i < n && i >= off) ? a[i - off] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = threadIdx.x + r * kBlock;
      if (i < n) a[i] += v[r];
    }
    __syncthreads();
  }
}

// Pairs among the moved beads (one molecule): new-new minus old-old with the
// reference's own pair functions (BBDist LJ, image-loop erfc).
__device__ void mm_block(const Geo& g, const Dev& d, int blk, int nblk, double* sh) {
  const int nm = d.rec->n_moved;
  const long long npair = (long long)nm * (nm - 1) / 2;
  double elj = 0, ere = 0;
  for (long long pidx = blk * (long long)kBlock + threadIdx.x; pidx < npair;
       pidx += (long long)nblk * kBlock) {
    const double tn = 2.0 * nm - 1;
    long long a = (long long)floor((tn - sqrt(tn * tn - 8.0 * (double)pidx)) / 2);
    if (a < 0) a = 0;
    if (a > nm - 2) a = nm - 2;
    while (a > 0 && a * (2LL * nm - a - 1) / 2 > pidx) a--;
    while (a < nm - 2 && (a + 1) * (2LL * nm - a - 2) / 2 <= pidx) a++;
    long long b = pidx - a * (2LL * nm - a - 1) / 2 + a + 1;
    if (b >= nm) b = nm - 1;
    const double4 na4 = d.mv_new[a], nb4 = d.mv_new[b], oa4 = d.mv_old[a], ob4 = d.mv_old[b];
    const double na[3] = {na4.x, na4.y, na4.z}, nb[3] = {nb4.x, nb4.y, nb4.z};
    const double oa[3] = {oa4.x, oa4.y, oa4.z}, ob[3] = {ob4.x, ob4.y, ob4.z};
    if (g.use_pair) {
      const LJPair lp = d.lj[d.mv_type[a] * g.n_types + d.mv_type[b]];
      elj += lj_energy(bb_dist(na, nb, g.Lsim, g.npbc), lp) - lj_energy(bb_dist(oa, ob, g.Lsim, g.npbc), lp);
    }
    if (g.use_ewald) ere += real_pair(na, na4.w, nb, nb4.w, g) - real_pair(oa, oa4.w, ob, ob4.w, g);
  }
  const double a = block_sum<kBlock>(elj, sh);
  const double c = block_sum<kBlock>(ere, sh);
  if (threadIdx.x == 0) d.part_mm[blk] = make_double2(a, c);
}

// Real-space (LJ + Ewald erfc) energy of one moved bead at its old (sign -1)
// or new (sign +1) position against every unmoved bead image within the
// search radius.  Blocks [0, nreal) = (moved bead, old/new, part): each query
// is split over `nsplit` blocks that take interleaved candidate slots;
// blocks [nreal, nreal + nmm) = moved-moved pairs.  The visited cells are
// listed in LDS with a prefix over their counts, so the candidate slots are
// spread evenly; the moved beads are skipped by index range when they are
// contiguous (every reference move generator moves a contiguous range).
template <int P>
struct RSmem {
  int s_cell[kMaxQueryCells];
  int s_img[kMaxQueryCells];
  int s_end[kMaxQueryCells];
  float s_rel[P ? kMaxQueryCells : 1][3];
  double sh[kBlock / kWave];
};

// LJ + real-space energy of one point p (charge p.w, type tm) against every
// bead image in the cell list within the search radius, skipping the index
// range [mlo, mhi] (contig) or the moved flags; candidate slots split over
// `nsplit` blocks (this block takes share `part`).  Result valid in thread 0.
template <int P>
__device__ double2 real_query(const Geo& g, const Dev& d, double4 p, int tm, int contig, int mlo, int mhi, int part,
                              int nsplit, RSmem<P>& sm) {
  int* s_cell = sm.s_cell;
  int* s_img = sm.s_img;
  int* s_end = sm.s_end;
  float (&s_rel)[P ? kMaxQueryCells : 1][3] = sm.s_rel;
  double* sh = sm.sh;
  const double pw[3] = {wrap_coord(p.x, g.L[0], g.inv_L[0]), wrap_coord(p.y, g.L[1], g.inv_L[1]),
                        wrap_coord(p.z, g.L[2], g.inv_L[2])};
  int lo[3], nd[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)floor((pw[k] - g.rs) * g.inv_h[k]);
    const int hi = (int)floor((pw[k] + g.rs) * g.inv_h[k]);
    nd[k] = hi - lo[k] + 1;
  }
  const int ncq = nd[0] * nd[1] * nd[2];
  double elj = 0, ere = 0;
  const LJPair* ljrow = d.lj + tm * g.n_types;
  const LJPair32* ljrow32 = d.lj32 + tm * g.n_types;
  const float qpf = (float)p.w;
  const float rs2f_wide = (float)(g.rs2 * (1.0 + 1e-4)) + 1e-4f;
  const float saf = (float)g.sqrt_alpha, lBf = (float)g.lB;
  const double lj_close = g.fp64_below;
  for (int base = 0; base < ncq; base += kMaxQueryCells) {
    const int nloc = min(kMaxQueryCells, ncq - base);
    for (int t = threadIdx.x; t < nloc; t += kBlock) {
      const int c = base + t;
      const int cz = c % nd[2], cy = (c / nd[2]) % nd[1], cx = c / (nd[2] * nd[1]);
      const int u[3] = {lo[0] + cx, lo[1] + cy, lo[2] + cz};
      int cell = 0, img = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int q = floordiv(u[k], g.nc[k]);
        const int w = u[k] - q * g.nc[k];
        cell = cell * g.nc[k] + w;
        img |= ((q + 128) & 0xff) << (8 * k);
        if constexpr (P != 0) s_rel[t][k] = (float)(pw[k] - (double)u[k] * g.h[k]);
      }
      s_cell[t] = cell;
      s_img[t] = img;
      // cells whose box (widened by a rounding margin) lies beyond the search
      // radius hold no candidate
      double dmin2 = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double b0 = (double)u[k] * g.h[k] - 1e-9 * g.L[k], b1 = b0 + g.h[k] + 2e-9 * g.L[k];
        const double dd = fmax(0.0, fmax(b0 - pw[k], pw[k] - b1));
        dmin2 += dd * dd;
      }
      s_end[t] = dmin2 > g.rs2 ? 0 : d.ccount[cell];
    }
    __syncthreads();
    lds_scan(s_end, nloc);
    const int total = s_end[nloc - 1];
    for (int it = part * kBlock + threadIdx.x; it < total; it += nsplit * kBlock) {
      int lo_i = 0, hi_i = nloc - 1;
      while (lo_i < hi_i) {
        const int mid = (lo_i + hi_i) >> 1;
        if (s_end[mid] > it) hi_i = mid; else lo_i = mid + 1;
      }
      const int t = lo_i;
      const int slot = it - (t ? s_end[t - 1] : 0);
      const size_t si = (size_t)s_cell[t] * g.cap + slot;
      const int j = d.cbead[si];
      if (j < 0 || (contig ? (j >= mlo && j <= mhi) : d.moved[j])) continue;
      const int iz = ((s_img[t] >> 16) & 0xff) - 128;
      const bool lj_ok = g.use_pair && (g.npbc == 3 || iz == 0);
      if constexpr (P == 0) {
        const double4 pj = d.cpos[si];
        const int img = s_img[t];
        const int ix = (img & 0xff) - 128, iy = ((img >> 8) & 0xff) - 128;
        const double dx = pj.x + ix * g.L[0] - pw[0];
        const double dy = pj.y + iy * g.L[1] - pw[1];
        const double dz = pj.z + iz * g.L[2] - pw[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > g.rs2) continue;
        const double r = sqrt(r2);
        if (lj_ok) elj += lj_energy(r, ljrow[d.ctype[si]]);
        if (g.use_ewald) {
          const double qq = p.w * pj.w;
          if (qq != 0 && r > 0 && r <= g.rc) ere += g.lB * qq * erfc_nodes(g, g.sqrt_alpha * r) / r;
        }
      } else {
        // FP32 prefilter on cell-relative coordinates (slightly widened so
        // no pair the FP64 test keeps is dropped), then the exact FP64
        // distance and cutoff test, then FP32 pair math.
        const float4 pr = d.cpos32[si];
        const float fx = pr.x - s_rel[t][0], fy = pr.y - s_rel[t][1], fz = pr.z - s_rel[t][2];
        if (fx * fx + fy * fy + fz * fz > rs2f_wide) continue;
        const double4 pj = d.cpos[si];
        const int img = s_img[t];
        const int ix = (img & 0xff) - 128, iy = ((img >> 8) & 0xff) - 128;
        const double dx = pj.x + ix * g.L[0] - pw[0];
        const double dy = pj.y + iy * g.L[1] - pw[1];
        const double dz = pj.z + iz * g.L[2] - pw[2];
        const double r2d = dx * dx + dy * dy + dz * dz;
        if (r2d > g.rs2) continue;
        const double rd = sqrt(r2d);
        if (rd < lj_close) {
          // close contacts carry the large, steep terms: FP64 like P == 0
          if (lj_ok) elj += lj_energy(rd, ljrow[d.ctype[si]]);
          if (g.use_ewald) {
            const double qq = p.w * pj.w;
            if (qq != 0 && rd > 0 && rd <= g.rc) ere += g.lB * qq * erfc_nodes(g, g.sqrt_alpha * rd) / rd;
          }
          continue;
        }
        const float r2 = (float)r2d;
        if (lj_ok) {
          const LJPair32 lp = ljrow32[d.ctype[si]];
          if (rd < lp.cut) {
            const float s2 = lp.sigma2 / r2;
            const float r6 = s2 * s2 * s2;
            elj += (double)(lp.eps4 * (r6 * r6 - r6) - lp.shift);
          }
        }
        if (g.use_ewald) {
          const float qq = qpf * pr.w;
          if (qq != 0.f && rd <= g.rc) {
            const float r = (float)rd;
            ere += (double)(lBf * qq * erfcf(saf * r) / r);
          }
        }
      }
    }
    __syncthreads();
  }
  const double a = block_sum<kBlock>(elj, sh);
  const double c = block_sum<kBlock>(ere, sh);
  return make_double2(a, c);
}

template <int P>
__device__ void real_block(const Geo& g, const Dev& d, int b, int nreal, int nmm, int nsplit, RSmem<P>& sm) {
  if (b >= nreal) {
    mm_block(g, d, b - nreal, nmm, sm.sh);
    return;
  }
  const int q2 = b / nsplit, part = b - q2 * nsplit;
  const int m = q2 >> 1, on = q2 & 1;
  const int nm = d.rec->n_moved;
  // one store for both cases (a separate zero store costs every wave a
  // register spill of the hoisted zero vector)
  double2 e = make_double2(0, 0);
  double sgn = 0;
  if (m < nm) {
    const double4 p = on ? d.mv_new[m] : d.mv_old[m];
    e = real_query<P>(g, d, p, d.mv_type[m], d.rec->contig, d.rec->lo, d.rec->hi, part, nsplit, sm);
    sgn = on ? 1.0 : -1.0;
  }
  if (threadIdx.x == 0) d.part_r[b] = make_double2(sgn * e.x, sgn * e.y);
}

// All dE work of one trial move in one launch: blocks [0, nkb) k space,
// then nreal real-space query blocks, then nmm moved-moved pair blocks.
template <int P>
__global__ __launch_bounds__(kBlock) void k_compute(Geo g, Dev d, int nkb, int nreal, int nmm, int nsplit) {
  __shared__ union U {
    KSmem k;
    RSmem<P> r;
  } sm;
  const int b = blockIdx.x;
  if (b < nkb) kspace_block(g, d, b, sm.k);
  else real_block<P>(g, d, b - nkb, nreal, nmm, nsplit, sm.r);
}

// ------------------------------------------------------------ epilogue
// Spring energy of molecule `mol` (potential_spring.cc:22-47: consecutive
// beads, no PBC).  Per-bond terms are summed by a fixed-order block tree, so
// the same bits come out whichever kernel evaluates the molecule.
__device__ double molecule_spring(const Geo& g, const Dev& d, int mol, bool trial, double* sbuf) {
  const int a = d.mol_start[mol], b = d.mol_start[mol + 1];
  double acc = 0;
  for (int base = a; base < b - 1; base += kBlock) {
    const int i = base + threadIdx.x;
    double e = 0;
    if (i < b - 1) {
      double p0[3] = {d.x[i], d.y[i], d.z[i]};
      double p1[3] = {d.x[i + 1], d.y[i + 1], d.z[i + 1]};
      if (trial) {
        const int s0 = d.mv_slot[i], s1 = d.mv_slot[i + 1];
        if (s0 >= 0) { const double4 v = d.mv_new[s0]; p0[0] = v.x; p0[1] = v.y; p0[2] = v.z; }
        if (s1 >= 0) { const double4 v = d.mv_new[s1]; p1[0] = v.x; p1[1] = v.y; p1[2] = v.z; }
      }
      const double dx = p0[0] - p1[0], dy = p0[1] - p1[1], dz = p0[2] - p1[2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      e = 0.5 * g.k_bond * (r - g.r0) * (r - g.r0);
    }
    acc += block_sum<kBlock>(e, sbuf);
  }
  return acc;  // valid in thread 0
}

// Spring energy of molecule `mol` with the beads lo..hi at mv_new[i - lo]
// (a proposal's contiguous moved range) and the others at their current
// positions; same fixed-order block tree as molecule_spring.
__device__ double molecule_spring_range(const Geo& g, const Dev& d, int mol, int lo, int hi,
                                        const double4* mv_new, double* sbuf) {
  const int a = d.mol_start[mol], b = d.mol_start[mol + 1];
  double acc = 0;
  for (int base = a; base < b - 1; base += kBlock) {
    const int i = base + threadIdx.x;
    double e = 0;
    if (i < b - 1) {
      double p0[3] = {d.x[i], d.y[i], d.z[i]};
      double p1[3] = {d.x[i + 1], d.y[i + 1], d.z[i + 1]};
      if (i >= lo && i <= hi) { const double4 v = mv_new[i - lo]; p0[0] = v.x; p0[1] = v.y; p0[2] = v.z; }
      if (i + 1 >= lo && i + 1 <= hi) { const double4 v = mv_new[i + 1 - lo]; p1[0] = v.x; p1[1] = v.y; p1[2] = v.z; }
      const double dx = p0[0] - p1[0], dy = p0[1] - p1[1], dz = p0[2] - p1[2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      e = 0.5 * g.k_bond * (r - g.r0) * (r - g.r0);
    }
    acc += block_sum<kBlock>(e, sbuf);
  }
  return acc;  // valid in thread 0
}

// Removes the tombstones of one cell, keeping slot order (one thread).
__device__ void compact_cell(const Geo& g, const Dev& d, int cell) {
  const int cnt = d.ccount[cell];
  int w = 0;
  for (int s = 0; s < cnt; s++) {
    const size_t si = (size_t)cell * g.cap + s;
    const int j = d.cbead[si];
    if (j < 0) continue;
    if (w != s) {
      const size_t sw = (size_t)cell * g.cap + w;
      d.cbead[sw] = j;
      d.cpos[sw] = d.cpos[si];
      d.cpos32[sw] = d.cpos32[si];
      d.ctype[sw] = d.ctype[si];
      d.bslot[j] = w;
    }
    w++;
  }
  for (int s = w; s < cnt; s++) d.cbead[(size_t)cell * g.cap + s] = -1;
  d.ccount[cell] = w;
}

// The moved beads of one commit: item k's trial position and bead index.
struct MoveItems {
  const double4* mv_new;
  const int* mv_idx;
  int n;
  __device__ __forceinline__ void get(int k, double4& p, int& i) const {
    p = mv_new[k];
    i = mv_idx[k];
  }
};

// Per-item scratch of a commit: new cell, key (the new cell of a bead that
// changes cell, -1 otherwise; read four at a time), arrival rank/count, and
// what the commit reads of the bead (cell, slot, type) and of its new cell
// (count), all read in one pass.
struct CommitSmem {
  alignas(16) int key[kMaxChainDev];
  int nc[kMaxChainDev];
  int rank[kMaxChainDev];
  int oc[kMaxChainDev];    // cell of the bead before the commit
  int os[kMaxChainDev];    // its slot there
  int ty[kMaxChainDev];    // its type
  int cnt[kMaxChainDev];   // count of its new cell
  int compacted;           // a destination cell was compacted (slots and counts moved)
};

// Positions and cell list of accepted moved beads (parallel, deterministic:
// a bead leaving its cell leaves a tombstone; arrivals in one cell take
// consecutive slots in item order).  Runs inside one block.  The items may
// span several accepted moves of distinct molecules: the beads left live in
// every cell, and their order, are those of committing the moves one after
// the other.  Everything the commit reads is read in one pass first; only a
// compaction (rare) moves slots and counts, and then they are read again.
// No barrier at the end: a caller that commits again, or reads what the
// commit wrote, synchronizes the block first.
template <class Items>
__device__ void commit_cells(const Geo& g, const Dev& d, const Items& items, CommitSmem& cs) {
  Ctl* C = d.ctl;
  const int nm = items.n;
  const int n4 = (nm + 3) >> 2;
  if (threadIdx.x == 0) cs.compacted = 0;
  // item threadIdx.x kept in registers when every item has its own thread
  double4 pn0 = make_double4(0, 0, 0, 0);
  int i0 = -1;
  // new cell of every moved bead
  for (int k = threadIdx.x; k < 4 * n4; k += kBlock) {
    if (k >= nm) {
      cs.key[k] = -1;
      continue;
    }
    double4 pn;
    int i;
    items.get(k, pn, i);
    if (k == threadIdx.x) {
      pn0 = pn;
      i0 = i;
    }
    const double xw = wrap_coord(pn.x, g.L[0], g.inv_L[0]);
    const double yw = wrap_coord(pn.y, g.L[1], g.inv_L[1]);
    const double zw = wrap_coord(pn.z, g.L[2], g.inv_L[2]);
    const int nc = (cell_coord(xw, g.inv_h[0], g.nc[0]) * g.nc[1] + cell_coord(yw, g.inv_h[1], g.nc[1])) *
                       g.nc[2] + cell_coord(zw, g.inv_h[2], g.nc[2]);
    const int oc = d.bcell[i], os = d.bslot[i], ty = d.type[i], cnt = d.ccount[nc];
    cs.nc[k] = nc;
    cs.key[k] = (oc != nc) ? nc : -1;
    cs.oc[k] = oc;
    cs.os[k] = os;
    cs.ty[k] = ty;
    cs.cnt[k] = cnt;
  }
  __syncthreads();
  // rank of each arrival among the arrivals in its cell (item order) and
  // their count, in one pass over the keys; a destination cell whose
  // high-water mark would pass the capacity is compacted first (by the
  // thread of its first arrival)
  const int4* k4 = reinterpret_cast<const int4*>(cs.key);
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    const int key = cs.key[k];
    if (key < 0) continue;
    int before = 0, total = 0;
#pragma unroll 4
    for (int q = 0; q < n4; q++) {
      const int4 v = k4[q];
      const int b = 4 * q;
      const int m0 = v.x == key, m1 = v.y == key, m2 = v.z == key, m3 = v.w == key;
      total += m0 + m1 + m2 + m3;
      before += (m0 & (b < k)) + (m1 & (b + 1 < k)) + (m2 & (b + 2 < k)) + (m3 & (b + 3 < k));
    }
    cs.rank[k] = before | (total << 16);
    if (before == 0 && cs.cnt[k] + total > g.cap) {
      compact_cell(g, d, key);
      cs.compacted = 1;
    }
  }
  __syncthreads();
  const bool reread = cs.compacted != 0;
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    double4 pn;
    int i;
    if (nm <= kBlock) {
      pn = pn0;
      i = i0;
    } else {
      items.get(k, pn, i);
    }
    d.x[i] = pn.x; d.y[i] = pn.y; d.z[i] = pn.z;
    const double xw = wrap_coord(pn.x, g.L[0], g.inv_L[0]);
    const double yw = wrap_coord(pn.y, g.L[1], g.inv_L[1]);
    const double zw = wrap_coord(pn.z, g.L[2], g.inv_L[2]);
    const int nc = cs.nc[k];
    const int cx = nc / (g.nc[1] * g.nc[2]), cy = (nc / g.nc[2]) % g.nc[1], cz = nc % g.nc[2];
    const float4 rel = make_float4((float)(xw - cx * g.h[0]), (float)(yw - cy * g.h[1]),
                                   (float)(zw - cz * g.h[2]), (float)pn.w);
    const int os = reread ? d.bslot[i] : cs.os[k];
    if (cs.key[k] < 0) {
      const size_t si = (size_t)nc * g.cap + os;
      d.cpos[si] = make_double4(xw, yw, zw, pn.w);
      d.cpos32[si] = rel;
      continue;
    }
    const int oc = cs.oc[k];
    if (oc >= 0) d.cbead[(size_t)oc * g.cap + os] = -1;
    const int rank = cs.rank[k] & 0xffff, total = cs.rank[k] >> 16;
    const int slot = (reread ? d.ccount[nc] : cs.cnt[k]) + rank;
    if (slot < g.cap) {
      const size_t sn = (size_t)nc * g.cap + slot;
      d.cbead[sn] = i;
      d.cpos[sn] = make_double4(xw, yw, zw, pn.w);
      d.cpos32[sn] = rel;
      d.ctype[sn] = cs.ty[k];
      d.bcell[i] = nc;
      d.bslot[i] = slot;
    } else {
      C->overflow = 1;
      d.bcell[i] = -1;
      d.bslot[i] = -1;
    }
    // the last arrival writes the cell's new count (after every thread has
    // read the counts when they are read again)
    const int cnt_new = rank == total - 1 ? min(slot + 1, g.cap) : -1;
    if (!reread && cnt_new >= 0) d.ccount[nc] = cnt_new;
    cs.rank[k] = cnt_new;
  }
  if (!reread) return;
  __syncthreads();
  for (int k = threadIdx.x; k < nm; k += kBlock)
    if (cs.key[k] >= 0 && cs.rank[k] >= 0) d.ccount[cs.nc[k]] = cs.rank[k];
}

__device__ void commit_block(const Geo& g, const Dev& d, int acc, CommitSmem& cs) {
  Ctl* C = d.ctl;
  const int nm = d.rec->n_moved;
  const int mol = d.rec->mol;
  if (acc) commit_cells(g, d, MoveItems{d.mv_new, d.mv_idx, nm}, cs);
  if (threadIdx.x == 0) {
    if (acc) {
      C->Mz += C->dMz;
      if (g.use_bond && d.mol_start[mol + 1] - d.mol_start[mol] > 1) d.bond_cur[mol] = C->bond_trial;
      if (g.dipole) C->cur_dipl = C->trial_dipl;
      for (int k = 0; k < 8; k++) C->running[k] += C->terms[k];
    } else if (g.dipole) {
      C->trial_dipl = C->cur_dipl;
    }
    if (d.rec->valid) {
      const int mt = d.rec->move_type;
      if (mt >= 0 && mt < 5) {
        C->attempted[mt] += 1;
        if (acc) C->accepted[mt] += 1;
      }
      C->valid_moves += 1;
    }
    C->pending = 0;
    C->move_counter += 1;
  }
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    const int i = d.mv_idx[k];
    d.moved[i] = 0;
    d.mv_slot[i] = -1;
  }
  __syncthreads();
}

// Reduction of all partials, bonded / wall / dipole terms and the
// ForceField::EnergyDifference combination with its early exits
// (force_field.cc:407-434) for the caller's move (host path); the caller
// decides and calls k_commit.
struct EpSmem {
  double sh[kBlock / kWave];
  double sbuf[kBlock];
  int s_forbid;
};

__device__ void epilogue_block(const Geo& g, const Dev& d, int nkb, int nrb, int nmmb, EpSmem& sm) {
  double* sh = sm.sh;
  double* sbuf = sm.sbuf;
  int& s_forbid = sm.s_forbid;
  Ctl* C = d.ctl;
  const int nm = d.rec->n_moved;
  const int mol = d.rec->mol;
  const int valid = d.rec->valid;
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
      // potential_external.cc:108-111: a forbidden bead makes the term 1e8
      t[T_WALL] = s_forbid ? kVeryLargeEnergy : ewt;
      dE += t[T_WALL];
      if (dE >= kVeryLargeEnergy) early = true;
    }
    if (!early && g.use_ewald) {
      t[T_REAL] = dr;
      t[T_KSPACE] = dk;
      double dd = 0;
      if (g.dipole) {
        if (g.dipole_exact) {
          const double mz1 = C->Mz + dmzt;
          C->trial_dipl = dipole_energy(g.dip_coef, mz1);
          dd = dipole_change(C->trial_dipl, dipole_energy(g.dip_coef, C->Mz));
        } else {
          // potential_ewald.cc:528-531: dipole of the CURRENT coordinates
          C->trial_dipl = dipole_energy(g.dip_coef, C->Mz);
          dd = dipole_change(C->trial_dipl, C->cur_dipl);
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
    C->dMz = dmzt;
    for (int k = 0; k < 8; k++) C->terms[k] = t[k];
    C->total = dE;
    C->pending = 1;
    C->sk_pending = 0;
    C->accept = 0;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_epilogue(Geo g, Dev d, int nkb, int nrb, int nmmb) {
  __shared__ EpSmem sm;
  epilogue_block(g, d, nkb, nrb, nmmb, sm);
}

// Host path: commit / roll back after the caller's decision.
__global__ __launch_bounds__(kBlock) void k_commit(Geo g, Dev d, int accept) {
  __shared__ CommitSmem cs;
  commit_block(g, d, accept && d.rec->valid, cs);
}

// Removes tombstones cell by cell, keeping slot order.
__global__ void k_compact(Geo g, Dev d) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.ncell) return;
  compact_cell(g, d, cell);
}

// ------------------------------------------------------------ proposals
__device__ __forceinline__ uint4 rnd(const McDev& mc, unsigned long long ctr, uint32_t stream,
                                     uint32_t sub) {
  return Philox::gen(make_uint4((uint32_t)ctr, (uint32_t)(ctr >> 32), stream, sub),
                     make_uint2((uint32_t)mc.seed, (uint32_t)(mc.seed >> 32)));
}

// randSphere (misc.cc:95-109, Marsaglia 1972) on the Philox stream `stream`.
__device__ void sphere(const McDev& mc, unsigned long long ctr, uint32_t stream, double v[3]) {
  double r1 = 0, r2 = 0, s = 2;
  for (uint32_t sub = 0; sub < 32 && s > 1; sub++) {
    const uint4 r = rnd(mc, ctr, stream, sub);
    r1 = 1 - 2 * u01_32(r.x);
    r2 = 1 - 2 * u01_32(r.y);
    s = r1 * r1 + r2 * r2;
    if (s > 1) {
      r1 = 1 - 2 * u01_32(r.z);
      r2 = 1 - 2 * u01_32(r.w);
      s = r1 * r1 + r2 * r2;
    }
  }
  if (s > 1) { r1 = 0; r2 = 0; s = 0; }
  const double ranh = 2 * sqrt(1 - s);
  v[0] = r1 * ranh;
  v[1] = r2 * ranh;
  v[2] = 1 - 2 * s;
}

// Simulation::TranslationalMove's choice of move type and molecule
// (simulation.cc:241-266) from Philox stream 0 of move `ctr`.
struct MoveSel {
  int mol, mt, valid;
};

__device__ MoveSel select_move(const Dev& d, const McDev& mc, unsigned long long ctr) {
  MoveSel s;
  const uint4 r0 = rnd(mc, ctr, 0, 0);
  int which_chain = (int)floor(u01_32(r0.x) * mc.n_chains);
  int which_ion = (int)floor(u01_32(r0.y) * mc.n_ions);
  if (which_chain == mc.n_chains) which_chain--;
  if (which_ion == mc.n_ions) which_ion--;
  int mt = 0;
  const double rn = u01_32(r0.z);
  double cum = mc.cum_prob[0];
  while (cum < rn && mt < 4) { mt++; cum += mc.cum_prob[mt]; }
  int mol = -1, valid = 1;
  if (mt == 0 && mc.n_ions > 0) mol = d.ion_ids[which_ion];
  else if (mt > 0 && mc.n_chains > 0) mol = d.chain_ids[which_chain];
  else valid = 0;
  if (mol < 0) { mol = 0; valid = 0; }
  s.mol = mol;
  s.mt = mt;
  s.valid = valid;
  return s;
}

// One proposal of Simulation::TranslationalMove (simulation.cc:241-312) with
// the move generators of molecule.cc:103-312, drawing from Philox4x32-10
// (counter = move number, streams: 1 move parameters, 2+i bead i).  `d` is
// the view of one proposal slot (its rec / mv_* arrays); nothing shared per
// bead is written, so two proposals of one molecule can coexist in a batch.
struct PrSmem {
  alignas(16) double s_b[kMaxChainDev][3];
  int s_mol, s_type, s_valid, s_a, s_len;
  double s_par[4];
  int s_ipar[3];
};

__device__ void propose_block(const Geo& g, const Dev& d, const McDev& mc, unsigned long long ctr, MoveSel sel,
                              PrSmem& sm) {
  double (&s_b)[kMaxChainDev][3] = sm.s_b;
  int& s_mol = sm.s_mol;
  int& s_type = sm.s_type;
  int& s_valid = sm.s_valid;
  int& s_a = sm.s_a;
  int& s_len = sm.s_len;
  double* s_par = sm.s_par;
  int* s_ipar = sm.s_ipar;
  if (threadIdx.x == 0) {
    const uint4 r1 = rnd(mc, ctr, 1, 0);
    s_mol = sel.mol;
    s_type = sel.mt;
    s_valid = sel.valid;
    s_a = d.mol_start[sel.mol];
    s_len = d.mol_start[sel.mol + 1] - s_a;
    s_par[0] = u01_32(r1.x);
    s_par[1] = u01_32(r1.y);
    s_par[2] = u01_32(r1.z);
    s_par[3] = u01_32(r1.w);
    s_ipar[0] = (int)(r1.w & 7u);
  }
  __syncthreads();
  const int mol = s_mol, mt = s_type, a = s_a, len = s_len;
  int nmv = 0;
  if (!s_valid || len > kMaxChainDev) {
    if (threadIdx.x == 0) {
      d.rec->n_moved = 0;
      d.rec->mol = mol;
      d.rec->move_type = mt;
      d.rec->valid = 0;
      d.rec->contig = 1;
      d.rec->lo = 0;
      d.rec->hi = -1;
    }
    return;
  }
  if (mt == 0) {
    // Molecule::BeadTranslate: 3 move_size along a random direction
    if (threadIdx.x == 0) {
      double v[3];
      sphere(mc, ctr, 2, v);
      double vl = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (vl > 0) vl = 3 * mc.move_size / vl;
      d.mv_new[0] = make_double4(d.x[a] + vl * v[0], d.y[a] + vl * v[1], d.z[a] + vl * v[2], d.q[a]);
    }
    nmv = 1;
  } else if (mt == 1) {
    // Molecule::COMTranslate: each component U[0, move_size/2], random sign
    double v[3];
    for (int k = 0; k < 3; k++) {
      v[k] = 0.5 * mc.move_size * s_par[k];
      if (((s_ipar[0] >> k) & 1u) == 0) v[k] = -v[k];
    }
    for (int i = threadIdx.x; i < len; i += kBlock)
      d.mv_new[i] = make_double4(d.x[a + i] + v[0], d.y[a + i] + v[1], d.z[a + i] + v[2], d.q[a + i]);
    nmv = len;
  } else if (mt == 2) {
    // Molecule::Pivot: each bond beyond the pivot gets a random kick of size
    // move_size*U, is renormalised to the (varied) bond length, and the
    // chain is rebuilt outward from the pivot bead.
    int pivot = (int)floor(len * s_par[0]);
    if (pivot == len) pivot--;
    const double msr = mc.move_size * s_par[1];
    for (int i = threadIdx.x; i < len; i += kBlock) {
      if (i == pivot) continue;
      const int nb = i > pivot ? i - 1 : i + 1;
      double v[3];
      sphere(mc, ctr, 2 + (uint32_t)i, v);
      double bx = d.x[a + i] - d.x[a + nb] + msr * v[0];
      double by = d.y[a + i] - d.y[a + nb] + msr * v[1];
      double bz = d.z[a + i] - d.z[a + nb] + msr * v[2];
      double bl = mc.bond_len;
      if (mc.vary_bond) {
        const uint4 rb = rnd(mc, ctr, 2 + (uint32_t)i, 0xB0u);
        bl += (mc.bond_len / 5.0) * (u01_32(rb.x) - 0.5);
      }
      const double nrm = bl / sqrt(bx * bx + by * by + bz * bz);
      s_b[i][0] = nrm * bx;
      s_b[i][1] = nrm * by;
      s_b[i][2] = nrm * bz;
    }
    __syncthreads();
    // the two halves outward from the pivot, one lane each, in LDS (each
    // bond vector is replaced by the position it leads to), then written
    // out in parallel
    if (threadIdx.x < 2) {
      const int dir = threadIdx.x == 0 ? 1 : -1;
      double px = d.x[a + pivot], py = d.y[a + pivot], pz = d.z[a + pivot];
      if (threadIdx.x == 0) {
        s_b[pivot][0] = px;
        s_b[pivot][1] = py;
        s_b[pivot][2] = pz;
      }
      // eight bonds are read together, then summed in order: one LDS
      // round trip per eight beads instead of one per bead, and no branch
      // inside a group
      const int cnt = dir > 0 ? len - 1 - pivot : pivot;
      int j = 0;
      for (; j + 8 <= cnt; j += 8) {
        double b[8][3];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int i = pivot + dir * (j + u + 1);
          b[u][0] = s_b[i][0];
          b[u][1] = s_b[i][1];
          b[u][2] = s_b[i][2];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int i = pivot + dir * (j + u + 1);
          px += b[u][0]; py += b[u][1]; pz += b[u][2];
          s_b[i][0] = px;
          s_b[i][1] = py;
          s_b[i][2] = pz;
        }
      }
      for (; j < cnt; j++) {
        const int i = pivot + dir * (j + 1);
        px += s_b[i][0]; py += s_b[i][1]; pz += s_b[i][2];
        s_b[i][0] = px;
        s_b[i][1] = py;
        s_b[i][2] = pz;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += kBlock)
      d.mv_new[i] = make_double4(s_b[i][0], s_b[i][1], s_b[i][2], d.q[a + i]);
    nmv = len;
  } else if (mt == 3) {
    // Molecule::Crankshaft: rotate beads first+1..last-1 about first->last
    int first = (int)floor((len - 2) * s_par[0]);
    int last = (int)floor((len - first) * s_par[1]) + first;
    if (last > len - 1) last = len - 1;
    double ax[3] = {d.x[a + last] - d.x[a + first], d.y[a + last] - d.y[a + first],
                    d.z[a + last] - d.z[a + first]};
    const double nn = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
    if (nn > 0) {
      const double inv = sqrt(nn);
      ax[0] /= inv; ax[1] /= inv; ax[2] /= inv;
    }
    const double angle = mc.move_size * (s_par[2] * 2 * M_PI - M_PI);
    double sn, cs;
    sincos(angle, &sn, &cs);
    double R[3][3];
    for (int i2 = 0; i2 < 3; i2++)
      for (int j2 = 0; j2 < 3; j2++) R[i2][j2] = (1 - cs) * ax[i2] * ax[j2] + (i2 == j2 ? cs : 0);
    R[0][1] -= sn * ax[2]; R[1][0] += sn * ax[2];
    R[0][2] += sn * ax[1]; R[2][0] -= sn * ax[1];
    R[1][2] -= sn * ax[0]; R[2][1] += sn * ax[0];
    const double fx = d.x[a + first], fy = d.y[a + first], fz = d.z[a + first];
    nmv = last - first - 1;
    if (nmv < 0) nmv = 0;
    for (int k = threadIdx.x; k < nmv; k += kBlock) {
      const int i = first + 1 + k;
      const double v0 = d.x[a + i] - fx, v1 = d.y[a + i] - fy, v2 = d.z[a + i] - fz;
      d.mv_new[k] = make_double4(R[0][0] * v0 + R[0][1] * v1 + R[0][2] * v2 + fx,
                                 R[1][0] * v0 + R[1][1] * v1 + R[1][2] * v2 + fy,
                                 R[2][0] * v0 + R[2][1] * v1 + R[2][2] * v2 + fz, d.q[a + i]);
      d.mv_idx[k] = a + i;
    }
  } else {
    // Molecule::RandomReptation: slide one slot, new end at the bond length
    double bl = mc.bond_len;
    if (mc.vary_bond) bl += (mc.bond_len / 5.0) * (s_par[0] - 0.5);
    const int dir = (s_ipar[0] & 1u) ? 1 : -1;
    const int end = dir == 1 ? len - 1 : 0;
    for (int i = threadIdx.x; i < len; i += kBlock) {
      if (i == end) {
        double v[3];
        sphere(mc, ctr, 2, v);
        const double vl = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        d.mv_new[i] = make_double4(d.x[a + i] + bl * v[0] / vl, d.y[a + i] + bl * v[1] / vl,
                                   d.z[a + i] + bl * v[2] / vl, d.q[a + i]);
      } else {
        const int s = i + dir;
        d.mv_new[i] = make_double4(d.x[a + s], d.y[a + s], d.z[a + s], d.q[a + i]);
      }
    }
    nmv = len;
  }
  __syncthreads();
  // moved list (crankshaft filled mv_idx itself)
  for (int k = threadIdx.x; k < nmv; k += kBlock) {
    const int i = (mt == 3) ? d.mv_idx[k] : a + k;
    if (mt != 3) d.mv_idx[k] = i;
    d.mv_old[k] = make_double4(d.x[i], d.y[i], d.z[i], d.q[i]);
    d.mv_type[k] = d.type[i];
  }
  if (threadIdx.x == 0) {
    d.rec->n_moved = nmv;
    d.rec->mol = mol;
    d.rec->move_type = mt;
    d.rec->valid = 1;
    // every generator moves a contiguous index range
    const int lo = (mt == 3) ? a + (int)floor((len - 2) * s_par[0]) + 1 : a;
    d.rec->contig = 1;
    d.rec->lo = lo;
    d.rec->hi = lo + nmv - 1;
  }
}

}  // namespace plum

namespace plum {

// One dimension's segment of a phase-table row (t = row start) by recurrence:
// e^{i l theta} = e^{i (l-1) theta} e^{i theta}, l = 0..n_d, theta_d =
// (2 kPi / L_d) x_d on the wrapped coordinate.
__device__ __forceinline__ void tables_row(const Geo& g, int dim, double xv, double2* t) {
  const double xw = wrap_coord(xv, g.L[dim], g.inv_L[dim]);
  const double k1 = (double)2 * kPi / g.L[dim];
  double sn, cs;
  sincos(k1 * xw, &sn, &cs);
  const double2 base = make_double2(cs, sn);
  const int n = dim == 0 ? g.nkx : (dim == 1 ? g.nky : g.nkz);
  t += dim == 0 ? 0 : (dim == 1 ? g.nkx + 1 : g.nkx + 1 + g.nky + 1);
  double2 e = make_double2(1.0, 0.0);
  t[0] = e;
  for (int l = 1; l <= n; l++) {
    e = cmul(e, base);
    t[l] = e;
  }
}

// Rows [r0, r0 + nr) of the phase tables of the pending move (row 2m + on:
// bead m, old/new), one block.  One thread per (row, dimension) runs
// tables_row into an LDS stage of rows padded to an odd number of entries
// (ws), and the block copies the rows out with consecutive threads on
// consecutive entries (a thread writing its own row straight to memory
// scatters every store over 64 rows); with no stage (ws = 0) the rows go
// straight to memory.  nr * 3 <= blockDim.x.
__device__ void tables_chunk(const Geo& g, const Dev& d, int r0, int nr, double2* stage, int ws) {
  const int w = g.tab_w;
  const int t = threadIdx.x;
  if (t < nr * 3) {
    const int lr = t / 3, dim = t - lr * 3;
    const int row = r0 + lr;
    const double4 p = (row & 1) ? d.mv_new[row >> 1] : d.mv_old[row >> 1];
    tables_row(g, dim, dim == 0 ? p.x : (dim == 1 ? p.y : p.z),
               ws ? stage + (size_t)lr * ws : d.tab + (size_t)row * w);
  }
  if (!ws) return;
  __syncthreads();
  // row and column of entry i advance by blockDim.x = q w + c per step
  double2* dst = d.tab + (size_t)r0 * w;
  const int q = blockDim.x / w, c = blockDim.x - q * w;
  int r = t / w, col = t - r * w;
#pragma unroll 4
  for (int i = t; i < nr * w; i += blockDim.x) {
    dst[i] = stage[r * ws + col];
    r += q;
    col += c;
    if (col >= w) {
      col -= w;
      r++;
    }
  }
}

}  // namespace plum
