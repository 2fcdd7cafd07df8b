// Trial energies of configurational-bias chain growth: ForceField::BeadsEnergy
// (cbmc.cc:5-151) for K trials of one growth step in one pass.  Trial t is a
// chain bead b1 = tri[t] and, for charged chains, its counterion b2 =
// tri[K + t]; the chain grown so far is grown[0..cur) (beads) and
// grown[cur .. 2 cur) (their counterions); a deletion excludes the chain being
// deleted and its counterions (bead index range [ex_lo, ex_hi]) from the
// existing system.
//   part 1  LJ + real space of b1, b2 against the existing beads (cell list)
//   part 2  LJ (not with the bonded previous bead) + real space against the
//           grown beads, explicit pairs
//   part 3  LJ(b1, b2), self terms, half self-images, real(b1, b2)
//   part 4  the walls
//   k space of all of it: 2 q_b sum_k c_k Re(e^{ik.r_b} conj(S_x(k))) with
//           S_x = S - S_excluded + S_grown, the (b1, b2) pair and the
//           self-images q_b^2 sum_k c_k
#pragma once
#include "kernels_move.h"

namespace plum {

constexpr int kCbmcMaxTrials = 1024;

struct CbmcIn {
  int K, cur, ex_lo, ex_hi;
  int type_c, type_i, charged, pad;
  double q_c, q_i;
};

// Parts 1: blocks (trial t, bead b, split s): LJ + real over the cell list.
template <int P>
__global__ __launch_bounds__(kBlock) void k_cbmc_cells(Geo g, Dev d, const CbmcIn* __restrict__ in,
                                                       const double* __restrict__ xyz, int nsplit,
                                                       double2* __restrict__ part) {
  __shared__ RSmem<P> sm;
  const int K = in->K;
  const int q = blockIdx.x / nsplit, s = blockIdx.x - (blockIdx.x / nsplit) * nsplit;
  const int t = q % K, b = q / K;   // b = 0 chain bead, 1 counterion
  double2 e = make_double2(0, 0);
  if (b == 0 || in->charged) {
    const double* r = xyz + 3 * (size_t)(b * K + t);
    const double4 p = make_double4(r[0], r[1], r[2], b ? in->q_i : in->q_c);
    e = real_query<P>(g, d, p, b ? in->type_i : in->type_c, 1, in->ex_lo, in->ex_hi, s, nsplit, sm);
  }
  if (threadIdx.x == 0) part[blockIdx.x] = e;
}

// k space, one thread per half-space k vector; per-block partial sums of each
// trial's k-space energy.
__device__ __forceinline__ double2 phase(const Geo& g, int4 l, const double* r) {
  double th = 0;
  const int ll[3] = {l.x, l.y, l.z};
#pragma unroll
  for (int k = 0; k < 3; k++) th += ((double)(ll[k] * 2) * kPi / g.L[k]) * wrap_coord(r[k], g.L[k], g.inv_L[k]);
  double sn, cs;
  sincos(th, &sn, &cs);
  return make_double2(cs, sn);
}

__global__ __launch_bounds__(kBlock) void k_cbmc_kspace(Geo g, Dev d, const CbmcIn* __restrict__ in,
                                                        const double* __restrict__ xyz,
                                                        const double* __restrict__ grown,
                                                        double* __restrict__ part) {
  __shared__ double sh[kBlock / kWave];
  const int K = in->K, cur = in->cur;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  double2 S = make_double2(0, 0);
  double ck = 0;
  int4 l = make_int4(0, 0, 0, 0);
  if (k < g.n_kh) {
    l = d.kl[k];
    ck = d.ck[k];
    S = d.Sk[k];
    for (int j = in->ex_lo; j <= in->ex_hi; j++) {
      const double r[3] = {d.x[j], d.y[j], d.z[j]};
      const double2 e = phase(g, l, r);
      S.x -= d.q[j] * e.x;
      S.y -= d.q[j] * e.y;
    }
    for (int j = 0; j < cur * (in->charged ? 2 : 1); j++) {
      const double2 e = phase(g, l, grown + 3 * (size_t)j);
      const double qj = j < cur ? in->q_c : in->q_i;
      S.x += qj * e.x;
      S.y += qj * e.y;
    }
  }
  for (int t = 0; t < K; t++) {
    double e = 0;
    if (k < g.n_kh) {
      const double2 e1 = phase(g, l, xyz + 3 * (size_t)t);
      // pairs with everything present, and the bead's own images
      e = 2.0 * in->q_c * ck * (e1.x * S.x + e1.y * S.y) + in->q_c * in->q_c * ck;
      if (in->charged) {
        const double2 e2 = phase(g, l, xyz + 3 * (size_t)(K + t));
        e += 2.0 * in->q_i * ck * (e2.x * S.x + e2.y * S.y) + in->q_i * in->q_i * ck;
        // the (b1, b2) pair
        e += 2.0 * in->q_c * in->q_i * ck * (e1.x * e2.x + e1.y * e2.y);
      }
    }
    const double r = block_sum<kBlock>(e, sh);
    if (threadIdx.x == 0) part[(size_t)t * gridDim.x + blockIdx.x] = r;
  }
}

// Slab dipole of the system less the excluded beads, then the grown chain
// beads, then their counterions (the M_z of DipoleEDiff before the trial
// joins, potential_ewald_coul.cc:485-520), one block.
__global__ __launch_bounds__(kBlock) void k_cbmc_mz(Geo g, Dev d, const CbmcIn* __restrict__ in,
                                                    const double* __restrict__ grown, double* __restrict__ mz) {
  __shared__ double sh[kBlock / kWave];
  double m = 0;
  for (int j = threadIdx.x; j < g.n; j += kBlock)
    if (j < in->ex_lo || j > in->ex_hi) m += d.q[j] * d.z[j];
  m = block_sum<kBlock>(m, sh);
  if (threadIdx.x == 0) {
    const int cur = in->cur;
    for (int i = 0; i < cur; i++) m += in->q_c * grown[3 * (size_t)i + 2];
    if (in->charged)
      for (int i = 0; i < cur; i++) m += in->q_i * grown[3 * (size_t)(cur + i) + 2];
    *mz = m;
  }
}

// Parts 2 and 3 explicit and the final sum, one block per trial; `mz` (slab
// dipole only) is the M_z k_cbmc_mz left.
__global__ __launch_bounds__(kBlock) void k_cbmc_final(Geo g, Dev d, const CbmcIn* __restrict__ in,
                                                       const double* __restrict__ xyz,
                                                       const double* __restrict__ grown,
                                                       const double2* __restrict__ cells, int nsplit,
                                                       const double* __restrict__ kpart, int nkb,
                                                       const double* __restrict__ mz, double* __restrict__ out) {
  __shared__ double sh[kBlock / kWave];
  const int K = in->K, cur = in->cur, t = blockIdx.x;
  const int ch = in->charged;
  const double* r1 = xyz + 3 * (size_t)t;
  const double* r2 = xyz + 3 * (size_t)(K + t);
  const LJPair* lj = d.lj;
  const int tc = in->type_c, ti = in->type_i, nt = g.n_types;
  double elj = 0, ere = 0;
  // part 2: the grown chain beads (i) and their counterions (cur + i)
  for (int i = threadIdx.x; i < cur; i += kBlock) {
    const double* gc = grown + 3 * (size_t)i;
    if (g.use_pair) {
      // LJ skips the bonded previous bead (i = cur - 1) of the chain bead
      if (i < cur - 1) elj += lj_energy(bb_dist(r1, gc, g.Lsim, g.npbc), lj[tc * nt + tc]);
      if (ch) {
        const double* gi = grown + 3 * (size_t)(cur + i);
        elj += lj_energy(bb_dist(r1, gi, g.Lsim, g.npbc), lj[tc * nt + ti]);
        elj += lj_energy(bb_dist(r2, gc, g.Lsim, g.npbc), lj[ti * nt + tc]);
        elj += lj_energy(bb_dist(r2, gi, g.Lsim, g.npbc), lj[ti * nt + ti]);
      }
    }
    if (g.use_ewald) {
      ere += real_pair(r1, in->q_c, gc, in->q_c, g);
      if (ch) {
        const double* gi = grown + 3 * (size_t)(cur + i);
        ere += real_pair(r1, in->q_c, gi, in->q_i, g);
        ere += real_pair(r2, in->q_i, gc, in->q_c, g);
        ere += real_pair(r2, in->q_i, gi, in->q_i, g);
      }
    }
  }
  // part 1 (cell-list partials) and the k-space partials
  for (int i = threadIdx.x; i < nsplit; i += kBlock) {
    elj += cells[(size_t)t * nsplit + i].x;
    ere += cells[(size_t)t * nsplit + i].y;
    if (ch) {
      elj += cells[(size_t)(K + t) * nsplit + i].x;
      ere += cells[(size_t)(K + t) * nsplit + i].y;
    }
  }
  double ek = 0;
  for (int i = threadIdx.x; i < nkb; i += kBlock) ek += kpart[(size_t)t * nkb + i];
  const double a = block_sum<kBlock>(elj, sh);
  const double c = block_sum<kBlock>(ere, sh);
  const double kk = block_sum<kBlock>(ek, sh);
  if (threadIdx.x == 0) {
    double pair_e = a;
    if (g.use_pair && ch) pair_e += lj_energy(bb_dist(r1, r2, g.Lsim, g.npbc), lj[tc * nt + ti]);
    if (g.use_ext) {
      // part 4: the walls
      pair_e += wall_energy(r1[2], d.wall[tc], g);
      if (ch) pair_e += wall_energy(r2[2], d.wall[ti], g);
    }
    double ewald_e = 0;
    if (g.use_ewald && pair_e < kVeryLargeEnergy) {
      // real space (parts 1, 2), self energies, half self-images, (b1, b2)
      ewald_e = c + kk + g.self_coef * in->q_c * in->q_c + 0.5 * real_pair(r1, in->q_c, r1, in->q_c, g);
      if (ch)
        ewald_e += g.self_coef * in->q_i * in->q_i + real_pair(r1, in->q_c, r2, in->q_i, g) +
                   0.5 * real_pair(r2, in->q_i, r2, in->q_i, g);
      if (g.dipole) ewald_e += dipole_ediff(g.dip_coef, *mz, in->q_c * r1[2], ch ? in->q_i * r2[2] : 0.0);
    }
    out[t] = pair_e >= kVeryLargeEnergy ? kVeryLargeEnergy : pair_e + ewald_e;
  }
}

}  // namespace plum
