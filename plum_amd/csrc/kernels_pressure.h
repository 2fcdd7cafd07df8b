// Volume-perturbation pressure: the energy change of one stretch of the box
// along z by kDz (ForceField::CalcPressureVolScalingHSELSlit,
// pressure.cc:187-387).  Beads' z are scaled by 1 + kDz/Lz (with the spring
// potential; grafted "L" beads stay, "R" beads move by kDz) or each molecule
// is shifted by kDz com_z / Lz (without it).  The pair part runs over all
// bead pairs once, new minus old per pair as the reference does, so the
// difference keeps its precision; the k-space part is summed per k vector from
// the change of S(k) and of c_k (k_vp_dsk).
#pragma once
#include "kernels_move.h"

namespace plum {

constexpr double kDz = 0.00001;   // utilities/constants.h:55
constexpr int kVpChunk = 512;     // beads per S(k) partial

// Stretched z of every bead.  One thread per molecule: the centre of mass in
// the reference's bead order, then the molecule's beads.
__global__ void k_vp_scale(Geo g, Dev d, double* zs) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.n_mols) return;
  const int a = d.mol_start[m], b = d.mol_start[m + 1];
  const double Lz = g.Lsim[2];
  double com = 0;
  for (int j = a; j < b; j++) com += d.z[j];
  com /= (b - a);
  double dcz = kDz * (com / Lz);
  const int k0 = d.wall[d.type[a]].kind;
  if (k0 == 2) dcz = kDz;
  else if (k0 == 1) dcz = 0;
  for (int j = a; j < b; j++) {
    const int kind = d.wall[d.type[j]].kind;
    zs[j] = (g.use_bond && kind == 0) ? d.z[j] * (1.0 + kDz / Lz) : d.z[j] + dcz;
  }
}

// PotentialEwaldCoul::PairEnergyRealForP (potential_ewald_coul.cc:167-196):
// the image sum in the box stretched along z, minimum image taken in that
// box; the single-image shortcut where no other image can reach rc.
__device__ inline double real_pair_stretch(const double* a, double qa, const double* b, double qb,
                                           const Geo& g) {
  if (qa * qb == 0) return 0.0;
  const double Ls[3] = {g.L[0], g.L[1], g.L[2] + kDz};
  double d[3];
  min_image(a, b, Ls, g.npbc, d);
  const double pref = g.lB * qa * qb;
  double e = 0;
  if (g.single_image && (g.npbc == 3 || fabs(d[2]) + g.rc * (1 + 1e-9) < Ls[2])) {
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r > 0 && r <= g.rc) e += pref * erfc_nodes(g, g.sqrt_alpha * r) / r;
    return e;
  }
  for (int i = -g.real_cell[0]; i <= g.real_cell[0]; i++)
    for (int j = -g.real_cell[1]; j <= g.real_cell[1]; j++)
      for (int k = -g.real_cell[2]; k <= g.real_cell[2]; k++) {
        const double x = d[0] + i * Ls[0], y = d[1] + j * Ls[1], z = d[2] + k * Ls[2];
        const double r = sqrt(x * x + y * y + z * z);
        if (r > 0 && r <= g.rc) e += pref * erfc_nodes(g, g.sqrt_alpha * r) / r;
      }
  return e;
}

// Unordered bead pairs a <= b (bead order is molecule order, so these are the
// reference's (i, j) <= (k, l) pairs), grid-stride over the triangular index;
// Ewald real-space and LJ new minus old per pair.  Pairs of phantom surface
// sites count only from the first half of the site list (pressure.cc:271-272).
__global__ __launch_bounds__(kBlock) void k_vp_pairs(Geo g, Dev d, const double* __restrict__ zs, int phantom,
                                                     long long npair, double* part) {
  __shared__ double sh[kBlock / kWave];
  const int n = g.n;
  const double Ls[3] = {g.Lsim[0], g.Lsim[1], g.Lsim[2] + kDz};
  double del = 0, dlj = 0;
  for (long long p = blockIdx.x * (long long)kBlock + threadIdx.x; p < npair; p += (long long)gridDim.x * kBlock) {
    // row a holds the pairs (a, a..n-1): offset a n - a (a - 1) / 2
    const double tn = 2.0 * n + 1;
    long long a = (long long)floor((tn - sqrt(tn * tn - 8.0 * (double)p)) / 2);
    if (a < 0) a = 0;
    if (a > n - 1) a = n - 1;
    while (a > 0 && a * n - a * (a - 1) / 2 > p) a--;
    while (a < n - 1 && (a + 1) * n - (a + 1) * a / 2 <= p) a++;
    const long long b = a + (p - (a * n - a * (a - 1) / 2));
    const int ia = (int)a, ib = (int)b;
    const int ma = d.mol_of[ia], mb = d.mol_of[ib];
    if (!(ma >= phantom || (ma < phantom / 2 && mb >= phantom / 2))) continue;
    const double ca[3] = {d.x[ia], d.y[ia], d.z[ia]}, cb[3] = {d.x[ib], d.y[ib], d.z[ib]};
    const double ta[3] = {ca[0], ca[1], zs[ia]}, tb[3] = {cb[0], cb[1], zs[ib]};
    if (g.use_ewald) {
      const double qa = d.q[ia], qb = d.q[ib];
      double o = real_pair(ca, qa, cb, qb, g), w = real_pair_stretch(ta, qa, tb, qb, g);
      if (ia == ib) {
        o *= 0.5;
        w *= 0.5;
      }
      del += w - o;
    }
    if (g.use_pair && ma >= phantom && mb >= phantom && ia != ib) {
      const LJPair lp = d.lj[d.type[ia] * g.n_types + d.type[ib]];
      dlj += lj_energy(bb_dist(ta, tb, Ls, g.npbc), lp) - lj_energy(bb_dist(ca, cb, g.Lsim, g.npbc), lp);
    }
  }
  const double s1 = block_sum<kBlock>(del, sh);
  const double s2 = block_sum<kBlock>(dlj, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// k space of the stretch, per half-space l vector of the current or the
// stretched k list (their union): S_c(k) of the three bead sets the pair rule
// separates -- mobile beads, the first and the second half of the phantom
// sites -- and the change dS_c of each under the stretch, as sums of
// per-bead phase changes e^{i k'.r'} - e^{i k.r} = 2 i sin(d/2) e^{i (th + th')/2}
// with d = th' - th small (x and y phases are unchanged; z is wrapped in the
// current box and the same image is used in the stretched box), so dE_k
// keeps its precision instead of being a difference of two totals.
// Bead chunk blockIdx.y.
__global__ __launch_bounds__(kBlock) void k_vp_dsk(Geo g, Dev d, const int4* __restrict__ kl, int nk,
                                                   const double* __restrict__ zs, int phantom, double2* part) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nk) return;
  const int4 l = kl[k];
  const double kx = (double)(l.x * 2) * kPi / g.L[0];
  const double ky = (double)(l.y * 2) * kPi / g.L[1];
  const double kz = (double)(l.z * 2) * kPi / g.L[2];
  const double Ls = g.L[2] + kDz;
  const double kzs = (double)(l.z * 2) * kPi / Ls;
  const int j0 = blockIdx.y * kVpChunk, j1 = min(g.n, j0 + kVpChunk);
  double2 S[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  double2 D[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int j = j0; j < j1; j++) {
    const double qj = d.q[j];
    if (qj == 0) continue;
    const int m = d.mol_of[j];
    const int c = m >= phantom ? 0 : (m < phantom / 2 ? 1 : 2);
    const double xw = wrap_coord(d.x[j], g.L[0], g.inv_L[0]);
    const double yw = wrap_coord(d.y[j], g.L[1], g.inv_L[1]);
    const double nz = floor(d.z[j] * g.inv_L[2]);
    const double zw = d.z[j] - nz * g.L[2];
    const double zws = zs[j] - nz * Ls;
    const double th = kx * xw + ky * yw + kz * zw;
    const double dth = kzs * zws - kz * zw;
    double sn, cs, sh, chh;
    sincos(th, &sn, &cs);
    sincos(th + 0.5 * dth, &sh, &chh);
    const double s2 = 2.0 * sin(0.5 * dth);
    S[c].x += qj * cs;
    S[c].y += qj * sn;
    // 2 i sin(d/2) e^{i (th + d/2)}
    D[c].x -= qj * s2 * sh;
    D[c].y += qj * s2 * chh;
  }
  for (int c = 0; c < 3; c++) {
    part[((size_t)blockIdx.y * 6 + c) * nk + k] = S[c];
    part[((size_t)blockIdx.y * 6 + 3 + c) * nk + k] = D[c];
  }
}

__device__ __forceinline__ double re_conj(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }

// dE_k per l of the union list: c'(Q(S + dS)) - c Q(S) with
// Q(S) = |S_M|^2 + 2 Re(S_A* S_B) + 2 Re(S_A* S_M) (the included pairs),
// written as dc Q(S) + c' (Q(S + dS) - Q(S)) where l is in both lists
// (dc = c' - c, evaluated on the host without cancellation); block partials.
__global__ __launch_bounds__(kBlock) void k_vp_dsk_energy(const double2* __restrict__ part, int nchunk, int nk,
                                                          const double* __restrict__ c_old,
                                                          const double* __restrict__ c_new,
                                                          const double* __restrict__ dc, double* epart) {
  __shared__ double sh[kBlock / kWave];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  double e = 0;
  if (k < nk) {
    double2 S[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    double2 D[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int ch = 0; ch < nchunk; ch++)
      for (int c = 0; c < 3; c++) {
        const double2 v = part[((size_t)ch * 6 + c) * nk + k];
        const double2 w = part[((size_t)ch * 6 + 3 + c) * nk + k];
        S[c].x += v.x;
        S[c].y += v.y;
        D[c].x += w.x;
        D[c].y += w.y;
      }
    const double2 M = S[0], A = S[1], B = S[2], dM = D[0], dA = D[1], dB = D[2];
    const double q0 = re_conj(M, M) + 2.0 * re_conj(A, B) + 2.0 * re_conj(A, M);
    const double dq = 2.0 * re_conj(M, dM) + re_conj(dM, dM) +
                      2.0 * (re_conj(A, dB) + re_conj(dA, B) + re_conj(dA, dB)) +
                      2.0 * (re_conj(A, dM) + re_conj(dA, M) + re_conj(dA, dM));
    const double co = c_old[k], cn = c_new[k];
    if (co != 0 && cn != 0) e = dc[k] * q0 + cn * dq;
    else e = cn * (q0 + dq) - co * q0;
  }
  const double r = block_sum<kBlock>(e, sh);
  if (threadIdx.x == 0) epart[blockIdx.x] = r;
}

// Walls (mobile beads), springs and the slab dipole, then every partial into
// out[0..5] = Ewald pairs, LJ pairs, walls, springs, dipole, k space.
__global__ __launch_bounds__(kBlock) void k_vp_small(Geo g, Dev d, const double* __restrict__ zs, int phantom,
                                                     const double* __restrict__ ppart, int npb,
                                                     const double* __restrict__ ekpart, int nkb, double* out) {
  __shared__ double sh[kBlock / kWave];
  Geo gs = g;
  gs.Lsim[2] = g.Lsim[2] + kDz;
  double ew = 0, eb = 0, mo = 0, mn = 0;
  for (int j = threadIdx.x; j < g.n; j += kBlock) {
    if (g.use_ext && d.mol_of[j] >= phantom) {
      const WallType w = d.wall[d.type[j]];
      ew += wall_energy(zs[j], w, gs) - wall_energy(d.z[j], w, g);
    }
    mo += d.q[j] * d.z[j];
    mn += d.q[j] * zs[j];
  }
  if (g.use_bond) {
    for (int m = threadIdx.x; m < g.n_mols; m += kBlock) {
      const int a = d.mol_start[m], b = d.mol_start[m + 1];
      double e = 0;
      for (int i = a; i < b - 1; i++) {
        const double dx = d.x[i] - d.x[i + 1], dy = d.y[i] - d.y[i + 1], dz = zs[i] - zs[i + 1];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        e += 0.5 * g.k_bond * (r - g.r0) * (r - g.r0);
      }
      eb += e - d.bond_cur[m];
    }
  }
  double pe = 0, pl = 0, kk = 0;
  for (int b = threadIdx.x; b < npb; b += kBlock) {
    pe += ppart[2 * b];
    pl += ppart[2 * b + 1];
  }
  for (int b = threadIdx.x; b < nkb; b += kBlock) kk += ekpart[b];
  const double t_ew = block_sum<kBlock>(ew, sh);
  const double t_eb = block_sum<kBlock>(eb, sh);
  const double t_mo = block_sum<kBlock>(mo, sh);
  const double t_mn = block_sum<kBlock>(mn, sh);
  const double t_pe = block_sum<kBlock>(pe, sh);
  const double t_pl = block_sum<kBlock>(pl, sh);
  const double t_kk = block_sum<kBlock>(kk, sh);
  if (threadIdx.x == 0) {
    double dip = 0;
    if (g.use_ewald && g.dipole) {
      const double* L = g.Lsim;
      dip = (g.lB * 2 * kPi) * (t_mn * t_mn / (L[0] * L[1] * (L[2] + kDz)) - t_mo * t_mo / (L[0] * L[1] * L[2]));
    }
    out[0] = t_pe;
    out[1] = g.use_pair ? t_pl : 0.0;
    out[2] = g.use_ext ? t_ew : 0.0;
    out[3] = g.use_bond ? t_eb : 0.0;
    out[4] = dip;
    out[5] = g.use_ewald ? t_kk : 0.0;
  }
}

}  // namespace plum
