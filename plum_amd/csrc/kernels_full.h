// From-scratch energy kernels: real space + LJ over all pairs through the
// cell list, S(k) by direct summation, springs / wall / self / dipole, and the
// reduction into the control block.
#pragma once
#include "kernels_move.h"

namespace plum {

// ============================================================ from scratch
// Real space + LJ over all pairs: one thread per occupied cell slot; each
// pair (and each self image, weight 1/2) is counted from its lower index.
__global__ __launch_bounds__(kBlock) void k_full_real(Geo g, Dev d) {
  __shared__ double sh[kBlock / kWave];
  const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
  double elj = 0, ere = 0;
  if (gid < (long long)g.ncell * g.cap) {
    const int cell = (int)(gid / g.cap), slot = (int)(gid % g.cap);
    const size_t si = (size_t)cell * g.cap + slot;
    const int i = slot < d.ccount[cell] ? d.cbead[si] : -1;
    if (i >= 0) {
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
              if (j < i) continue;  // also skips tombstones (-1)
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
                  const double e = g.lB * qq * erfc_nodes(g, g.sqrt_alpha * r) / r;
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
      if (g.dipole) f[T_DIPOLE] = dipole_energy(g.dip_coef, mzt);
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

}  // namespace plum
