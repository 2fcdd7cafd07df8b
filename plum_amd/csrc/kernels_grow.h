// Configurational-bias growth of a whole chain on the device: every growth
// step of an insertion, a deletion or a Widom test insertion (cbmc.cc:213-533)
// runs as two launches with no host round trip in between, the trial draws
// made beforehand by the host in the reference's order.
//   k_grow_init    S_x(k) = S(k) - S_excluded(k), Rosenbluth weight 1
//   per step s:
//   k_grow_trials  part 1 (cell list, LJ + real space), parts 2-4 (the grown
//                  chain, the bead-counterion pair, self images, walls; one
//                  block per trial, threads over pairs x images) and the
//                  k-space sums of each trial bead and counterion, positions
//                  built on the device from the bead chosen at step s - 1
//   k_grow_select  the trial energies from those partial sums, the weights,
//                  W *= W_s / K, the choice of the next bead (insertion: by
//                  the pre-drawn uniform; deletion: the existing bead, trial
//                  0), and S_x += the chosen pair
// Step 0 is the first bead (and its counterion) alone, trial 0 of K.
#pragma once
#include "kernels_cbmc.h"

namespace plum {

struct GrowIn {
  int K, L, mode, charged;   // mode 0 insertion, 1 deletion, 2 test insertion
  int type_c, type_i, ex_lo, ex_hi;
  double q_c, q_i;
};

static_assert(sizeof(GrowIn) % alignof(double) == 0, "the random numbers follow GrowIn in one buffer");

struct GrowOut {
  double W;       // Rosenbluth weight
  int stop;       // step at which the weight reached 0 (-1: grown to the end)
  int pad;
  // slab dipole: sum of q z over the system less the excluded beads plus the
  // chain beads chosen so far, and over the counterions chosen so far
  double mz_c, mz_i;
};

// Random numbers and known positions of one growth, one upload:
//   first[6]            first chain bead, its counterion (insertion)
//   disp[L K 3]         step s, trial t: bond vector from the bead of step s-1
//   ion[L K 3]          step s, trial t: counterion position
//   sel[L]              step s: uniform of the choice among the trials
//   act[2 L 3]          deletion: the chain's beads, then its counterions
__host__ __device__ inline size_t grow_rnd_size(int K, int L) { return 6 + 6 * (size_t)L * K + L + 6 * (size_t)L; }

struct GrowPtr {
  const double *first, *disp, *ion, *sel, *act;
};

__device__ __forceinline__ GrowPtr grow_ptr(const double* rnd, int K, int L) {
  GrowPtr p;
  p.first = rnd;
  p.disp = rnd + 6;
  p.ion = p.disp + 3 * (size_t)L * K;
  p.sel = p.ion + 3 * (size_t)L * K;
  p.act = p.sel + L;
  return p;
}

// Trial t of step s: chain bead r1 and counterion r2.
__device__ __forceinline__ void grow_trial(const GrowIn& in, const GrowPtr& p, const double* grown, int s, int t,
                                           double r1[3], double r2[3]) {
  const int L = in.L, K = in.K;
  if (s == 0) {
    const double* a = in.mode == 1 ? p.act : p.first;
    const double* b = in.mode == 1 ? p.act + 3 * (size_t)L : p.first + 3;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      r1[k] = a[k];
      r2[k] = b[k];
    }
  } else if (in.mode == 1 && t == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      r1[k] = p.act[3 * (size_t)s + k];
      r2[k] = p.act[3 * (size_t)(L + s) + k];
    }
  } else {
    const double* d = p.disp + 3 * ((size_t)s * K + t);
    const double* e = p.ion + 3 * ((size_t)s * K + t);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      r1[k] = grown[3 * (size_t)(s - 1) + k] + d[k];
      r2[k] = e[k];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_grow_init(Geo g, Dev d, const GrowIn* __restrict__ in,
                                                      double2* __restrict__ Sx, GrowOut* __restrict__ out) {
  __shared__ double sh[kBlock / kWave];
  if (blockIdx.x == 0 && g.dipole && g.use_ewald) {
    // M_z of the system less the excluded beads (DipoleEDiff's molecule
    // loop, potential_ewald_coul.cc:485-506)
    double m = 0;
    for (int j = threadIdx.x; j < g.n; j += kBlock)
      if (j < in->ex_lo || j > in->ex_hi) m += d.q[j] * d.z[j];
    m = block_sum<kBlock>(m, sh);
    if (threadIdx.x == 0) {
      out->mz_c = m;
      out->mz_i = 0;
    }
  }
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (g.use_ewald && k < g.n_kh) {
    const int4 l = d.kl[k];
    double2 S = d.Sk[k];
    for (int j = in->ex_lo; j <= in->ex_hi; j++) {
      const double r[3] = {d.x[j], d.y[j], d.z[j]};
      const double2 e = phase(g, l, r);
      S.x -= d.q[j] * e.x;
      S.y -= d.q[j] * e.y;
    }
    Sx[k] = S;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out->W = 1.0;
    out->stop = -1;
  }
}

// One image (index img of the reference's image range) of the real-space
// pair sum of real_pair_ref.
__device__ __forceinline__ double real_pair_image(const double* a, double qa, const double* b, double qb,
                                                  const Geo& g, int img) {
  const int nj = 2 * g.real_cell[1] + 1, nk = 2 * g.real_cell[2] + 1;
  const int i = img / (nj * nk) - g.real_cell[0];
  const int j = (img / nk) % nj - g.real_cell[1];
  const int k = img % nk - g.real_cell[2];
  double dv[3];
  min_image(a, b, g.L, g.npbc, dv);
  const double x = dv[0] + i * g.L[0], y = dv[1] + j * g.L[1], z = dv[2] + k * g.L[2];
  const double r = sqrt(x * x + y * y + z * z);
  return (r > 0 && r <= g.rc) ? g.lB * qa * qb * erfc_nodes(g, g.sqrt_alpha * r) / r : 0.0;
}

// Parts 2-4 of trial t (cbmc.cc:52-151): LJ with the grown chain (not the
// bonded previous bead) and real space with it over every image, the
// bead-counterion pair, the half self images, the walls.  Returns (pair
// energy, real-space energy) in thread 0.
__device__ double2 grow_explicit(const Geo& g, const Dev& d, const GrowIn& in, const double* grown, int cur,
                                 const double r1[3], const double r2[3], double* sh) {
  const int L = in.L, ch = in.charged, tc = in.type_c, ti = in.type_i, nt = g.n_types;
  const LJPair* lj = d.lj;
  double elj = 0, ere = 0;
  if (g.use_pair)
    for (int q = threadIdx.x; q < 4 * cur; q += kBlock) {
      const int i = q % cur, kind = q / cur;   // (r1|r2) x (chain|ion)
      if (kind > 0 && !ch) continue;
      if (kind == 0 && i == cur - 1) continue;
      const double* a = kind < 2 ? r1 : r2;
      const double* b = grown + 3 * (size_t)((kind & 1) ? L + i : i);
      elj += lj_energy(bb_dist(a, b, g.Lsim, g.npbc), lj[((kind < 2) ? tc : ti) * nt + ((kind & 1) ? ti : tc)]);
    }
  if (g.use_ewald) {
    const int ni = (2 * g.real_cell[0] + 1) * (2 * g.real_cell[1] + 1) * (2 * g.real_cell[2] + 1);
    const int np = 4 * cur + 3;
    for (int it = threadIdx.x; it < np * ni; it += kBlock) {
      const int q = it / ni, img = it - q * ni;
      const double *a, *b;
      double qa, qb, f = 1.0;
      if (q < 4 * cur) {
        const int i = q % cur, kind = q / cur;
        if (kind > 0 && !ch) continue;
        a = kind < 2 ? r1 : r2;
        qa = kind < 2 ? in.q_c : in.q_i;
        b = grown + 3 * (size_t)((kind & 1) ? L + i : i);
        qb = (kind & 1) ? in.q_i : in.q_c;
      } else if (q == 4 * cur) {
        a = b = r1;
        qa = qb = in.q_c;
        f = 0.5;
      } else {
        if (!ch) continue;
        a = r2;
        qa = in.q_i;
        b = q == 4 * cur + 1 ? r2 : r1;
        qb = q == 4 * cur + 1 ? in.q_i : in.q_c;
        f = q == 4 * cur + 1 ? 0.5 : 1.0;
      }
      if (qa * qb != 0) ere += f * real_pair_image(a, qa, b, qb, g, img);
    }
  }
  const double a = block_sum<kBlock>(elj, sh);
  const double b = block_sum<kBlock>(ere, sh);
  double pair_e = a;
  if (threadIdx.x == 0) {
    if (g.use_pair && ch) pair_e += lj_energy(bb_dist(r1, r2, g.Lsim, g.npbc), lj[tc * nt + ti]);
    if (g.use_ext) {
      pair_e += wall_energy(r1[2], d.wall[tc], g);
      if (ch) pair_e += wall_energy(r2[2], d.wall[ti], g);
    }
  }
  return make_double2(pair_e, b);
}

// Blocks [0, 2 K nsplit): part 1 of (bead b, trial t, split); then K blocks
// of parts 2-4, one per trial; then K x nkb blocks of k space, (trial, 256
// vectors) each.  Step 0 evaluates trial 0 only.
template <int P>
__global__ __launch_bounds__(kBlock) void k_grow_trials(Geo g, Dev d, const GrowIn* __restrict__ inp,
                                                        const double* __restrict__ rnd,
                                                        const double* __restrict__ grown,
                                                        const double2* __restrict__ Sx,
                                                        const GrowOut* __restrict__ out, int s, int nsplit,
                                                        double2* __restrict__ cells, double2* __restrict__ expl,
                                                        double* __restrict__ kpart) {
  __shared__ RSmem<P> sm;
  __shared__ double sh[kBlock / kWave];
  if (out->stop >= 0) return;
  const GrowIn in = *inp;
  const GrowPtr p = grow_ptr(rnd, in.K, in.L);
  const int K = in.K, Kt = s == 0 ? 1 : K;
  const int ncell = 2 * K * nsplit;
  if ((int)blockIdx.x >= ncell && (int)blockIdx.x < ncell + K) {
    const int t = blockIdx.x - ncell;
    if (t >= Kt) return;
    double r1[3], r2[3];
    grow_trial(in, p, grown, s, t, r1, r2);
    const double2 e = grow_explicit(g, d, in, grown, s, r1, r2, sh);
    if (threadIdx.x == 0) expl[t] = e;
    return;
  }
  if ((int)blockIdx.x < ncell) {
    const int q = blockIdx.x / nsplit, sp = blockIdx.x - q * nsplit;
    const int t = q % K, b = q / K;
    double2 e = make_double2(0, 0);
    if (t < Kt && (b == 0 || in.charged)) {
      double r1[3], r2[3];
      grow_trial(in, p, grown, s, t, r1, r2);
      const double* r = b ? r2 : r1;
      const double4 pos = make_double4(r[0], r[1], r[2], b ? in.q_i : in.q_c);
      e = real_query<P>(g, d, pos, b ? in.type_i : in.type_c, 1, in.ex_lo, in.ex_hi, sp, nsplit, sm);
    }
    if (threadIdx.x == 0) cells[blockIdx.x] = e;
    return;
  }
  // k space: block (trial t, chunk kb of 256 vectors)
  const int nkb = (gridDim.x - ncell - K) / K;
  const int b = blockIdx.x - ncell - K, t = b / nkb, kb = b - t * nkb;
  if (t >= Kt) return;
  const int k = kb * kBlock + threadIdx.x;
  double e = 0;
  if (k < g.n_kh) {
    const int4 l = d.kl[k];
    const double ck = d.ck[k];
    const double2 S = Sx[k];
    double r1[3], r2[3];
    grow_trial(in, p, grown, s, t, r1, r2);
    const double2 e1 = phase(g, l, r1);
    e = 2.0 * in.q_c * ck * (e1.x * S.x + e1.y * S.y) + in.q_c * in.q_c * ck;
    if (in.charged) {
      const double2 e2 = phase(g, l, r2);
      e += 2.0 * in.q_i * ck * (e2.x * S.x + e2.y * S.y) + in.q_i * in.q_i * ck;
      e += 2.0 * in.q_c * in.q_i * ck * (e1.x * e2.x + e1.y * e2.y);
    }
  }
  const double r = block_sum<kBlock>(e, sh);
  if (threadIdx.x == 0) kpart[(size_t)t * nkb + kb] = r;
}

constexpr int kSelThreads = 1024;

// One block: each trial's energy from the partial sums (part 1 per split,
// parts 2-4, k space), the weights, the Rosenbluth update and the choice.
__global__ __launch_bounds__(kSelThreads) void k_grow_select(Geo g, Dev d, const GrowIn* __restrict__ inp,
                                                             const double* __restrict__ rnd,
                                                             double* __restrict__ grown, double2* __restrict__ Sx,
                                                             GrowOut* __restrict__ out, int s, int nsplit, int nkb,
                                                             const double2* __restrict__ cells,
                                                             const double2* __restrict__ expl,
                                                             const double* __restrict__ kpart) {
  __shared__ double s_e[kCbmcMaxTrials], s_w[kCbmcMaxTrials];
  __shared__ double s_pos[6];
  __shared__ int s_go;
  if (out->stop >= 0) return;
  const GrowIn in = *inp;
  const GrowPtr p = grow_ptr(rnd, in.K, in.L);
  const int K = in.K, L = in.L, Kt = s == 0 ? 1 : K, ch = in.charged;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave, nw = kSelThreads / kWave;
  for (int t = wid; t < Kt; t += nw) {
    double elj = 0, ere = 0, ek = 0;
    for (int i = lane; i < nsplit; i += kWave) {
      const double2 a = cells[(size_t)t * nsplit + i];
      elj += a.x;
      ere += a.y;
      if (ch) {
        const double2 b = cells[(size_t)(K + t) * nsplit + i];
        elj += b.x;
        ere += b.y;
      }
    }
    for (int i = lane; i < nkb; i += kWave) ek += kpart[(size_t)t * nkb + i];
    elj = wave_sum(elj);
    ere = wave_sum(ere);
    ek = wave_sum(ek);
    if (lane == 0) {
      const double2 x = expl[t];
      const double pair_e = elj + x.x;
      double ewald_e = 0;
      if (g.use_ewald && pair_e < kVeryLargeEnergy) {
        ewald_e = ere + x.y + ek + g.self_coef * in.q_c * in.q_c;
        if (ch) ewald_e += g.self_coef * in.q_i * in.q_i;
        if (g.dipole) {
          double r1[3], r2[3];
          grow_trial(in, p, grown, s, t, r1, r2);
          ewald_e += dipole_ediff(g.dip_coef, out->mz_c + out->mz_i, in.q_c * r1[2], ch ? in.q_i * r2[2] : 0.0);
        }
      }
      s_e[t] = pair_e >= kVeryLargeEnergy ? kVeryLargeEnergy : pair_e + ewald_e;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // weights in trial order, summed as the reference sums them
    double Wi = 0;
    for (int t = 0; t < Kt; t++) {
      const double w = exp(-g.beta * s_e[t]);
      s_w[t] = w;
      Wi += w;
    }
    double W = out->W;
    int go = 1, cb = 0;
    if (s == 0) {
      W *= s_w[0];
      if (in.mode == 0 && W <= 0) go = 0;
    } else {
      W *= Wi / K;
      if (in.mode != 1 && W <= 0) go = 0;
      if (go && in.mode != 1) {
        const double rn = p.sel[s] * Wi;
        double cum = s_w[0];
        while (cum < rn && cb < K - 1) {
          cb++;
          cum += s_w[cb];
        }
      }
    }
    out->W = W;
    if (!go) out->stop = s;
    s_go = go;
    if (go) {
      double r1[3], r2[3];
      grow_trial(in, p, grown, s, cb, r1, r2);
      for (int k = 0; k < 3; k++) {
        grown[3 * (size_t)s + k] = r1[k];
        grown[3 * (size_t)(L + s) + k] = r2[k];
        s_pos[k] = r1[k];
        s_pos[3 + k] = r2[k];
      }
      if (g.dipole && g.use_ewald) {
        out->mz_c += in.q_c * r1[2];
        if (ch) out->mz_i += in.q_i * r2[2];
      }
    }
  }
  __syncthreads();
  // the chosen bead and counterion join S_x for the next step
  if (!s_go || s == L - 1 || !g.use_ewald || g.n_kh == 0) return;
  const double r1[3] = {s_pos[0], s_pos[1], s_pos[2]};
  const double r2[3] = {s_pos[3], s_pos[4], s_pos[5]};
  for (int k = threadIdx.x; k < g.n_kh; k += kSelThreads) {
    const int4 l = d.kl[k];
    const double2 e1 = phase(g, l, r1);
    double2 S = Sx[k];
    S.x += in.q_c * e1.x;
    S.y += in.q_c * e1.y;
    if (ch) {
      const double2 e2 = phase(g, l, r2);
      S.x += in.q_i * e2.x;
      S.y += in.q_i * e2.y;
    }
    Sx[k] = S;
  }
}

}  // namespace plum
