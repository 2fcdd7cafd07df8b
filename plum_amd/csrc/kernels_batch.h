// Device-resident Monte Carlo over R independent systems (R = 1 for a single
// system), in speculative batches.
//
// A batch makes up to `pmax` proposals (screen slots) per system; the ones to
// be evaluated exactly get one of kMaxProp exact slots.  Without delayed
// acceptance every proposal is evaluated exactly (pmax <= kMaxProp).  With
// it, a surrogate screens up to kMaxScreen proposals and only those passing
// the first stage (plus, as training data, a few that did not) are evaluated.
//
// One batch = five launches (kernel boundaries order them; no grid-wide
// fences, which on a multi-XCD part would write back and invalidate L2),
// plus the surrogate's two between proposal and plan:
//   k_propose_spec   grid (pmax, R): proposal s of system r (move number
//                    counter + s), its spring / wall / dipole parts and
//                    bounding sphere (a molecule proposed twice is proposed
//                    from its current coordinates the second time).
//   k_plan_spec      grid (R, 1 + pmax chunks + fold chunks): first-stage
//                    decisions, exact slots, block counts; pairs of exact
//                    slots within interaction range get cross-term blocks;
//                    beside it, the proposals' phase tables in chunks of rows
//                    and the fold of the previous batch's accepted dS(k) into
//                    S(k).
//   k_compute_spec   all systems' k-space waves (dS_e(k) of every exact slot
//                    and its dE_k per k block), real-space query blocks
//                    (cluster queries by waves) and moved-moved blocks per
//                    slot, and cross blocks (LJ + real pairs between two
//                    slots' moved beads, old/new combinations), spread over
//                    one grid.
//   k_reduce_spec    one block per sum and system, fixed order; the k-space
//                    cross sums 2 c_k Re(dS_p* dS_q) formed from the dS rows.
//   k_resolve_spec   grid R: the reference's sequential Metropolis chain over
//                    the batch (simulation.cc:241-355), decided by one wave,
//                    each dE corrected by the cross terms of the proposals
//                    accepted before it, then the commits.
// Every proposal is evaluated against the state before the batch; accepting
// p changes q's dE by exactly the cross terms, so the chain is the sequential
// one (the random numbers depend only on the move number).  A second
// proposal of one molecule assumes the first is rejected; if it was
// accepted, the batch ends before the second.  A surrogate value
// has no such correction: the batch ends before the first proposal whose
// descriptors could see a bead accepted earlier in the batch, and that
// proposal is made again, from the same move number, in the next batch.
#pragma once
#include "kernels_move.h"

namespace plum {

constexpr int kMaxReplicas = 1024;
constexpr int kMaxSplit = 8;           // most blocks per real-space query
constexpr int kDefaultSplitDiv = 384;   // candidate slots per query share
constexpr int kKSpec = 32;             // k-vector runs per speculative k-space wave

// Per-system launch shape.
struct RepCfg {
  int nkb_s;    // k-space blocks of kKSpec runs (0 without Ewald)
  int nsplit;   // blocks per real-space query
  int ncell;    // cells
  int pmax;     // proposals per batch (<= kMaxProp, <= kMaxScreen with the surrogate)
  int maxm;     // moved-list capacity per proposal (longest molecule)
  int nmmb;     // moved-moved block capacity per exact slot
  int rstride;  // real-space partials per exact slot
  int xcap;     // cross-term partial capacity
  int da;       // delayed acceptance on
  int fill;     // exact slots given to first-stage rejections (training data)
  float r_d;    // descriptor radius of the surrogate
  int train;    // this system feeds the surrogate's training ring
  int nrun;     // k-vector runs (Dev::krun)
  int qfill;    // waves of k_compute_spec resident on the chip at once
};

// Real-space query blocks of a move of nm beads: one query per cluster of up
// to kCluster consecutive moved beads at their old and one at their new
// positions, qw waves per query (each takes a share of the candidate slots),
// kQWaves waves per block.
constexpr int kQWaves = kBlock / kWave;
constexpr int kMaxQWave = 4 * kMaxSplit;
// One wave per split share; a batch with few exact proposals (fewer than 16
// over all systems) spreads each query over four times as many waves, as its
// queries alone would leave most of the chip idle, unless the system's own
// query waves (its exact slots' clusters at old and new positions, times the
// shares) would then outnumber the waves the chip holds at once (fill).  The
// plan chooses it per batch from the exact slots it assigned (with delayed
// acceptance most screened proposals never reach the queries).
__host__ __device__ __forceinline__ int query_waves(int nsplit, int proposals, int clusters, int fill) {
  if (nsplit == 1 || proposals >= 16) return nsplit;
  return (long long)clusters * nsplit * kQWaves <= fill ? kQWaves * nsplit : nsplit;
}
constexpr int kCluster = 8;   // moved beads per real-space query (cluster)
__host__ __device__ __forceinline__ int query_clusters(int nm) { return (nm + kCluster - 1) / kCluster; }
__host__ __device__ __forceinline__ int query_blocks(int nm, int qw) {
  return (2 * query_clusters(nm) * qw + kQWaves - 1) / kQWaves;
}

// k-space blocks per k block: one wave per exact slot.
__host__ __device__ __forceinline__ int kspace_groups(int ne) { return (ne + kQWaves - 1) / kQWaves; }

// Moved-moved pair blocks: 256 pairs each, independent of the launch size.
__host__ __device__ __forceinline__ int nmm_blocks(int nm) {
  const int npair = nm * (nm - 1) / 2;
  return npair > 0 ? (npair + kBlock - 1) / kBlock : 1;
}

__host__ __device__ __forceinline__ int cross_blocks(int np, int nq) { return (np * nq + kBlock - 1) / kBlock; }

template <typename T>
__device__ __forceinline__ T load_const(const T* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(4))) const T CT;
  return ((CT*)p)[i];
#else
  return p[i];
#endif
}

__device__ __forceinline__ int uload(const int* p) { return __builtin_amdgcn_readfirstlane(*p); }

// A value every lane of the wave holds, kept in scalar registers.
__device__ __forceinline__ double uniform_double(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Set bits of m below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// System r owning virtual block v: first r with pre[r] > v (inclusive prefix).
__device__ __forceinline__ int find_replica(const int* pre, int R, int v) {
  int lo = 0, hi = R - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] > v) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// View of proposal s (move record, moved list, phase tables) with the dS(k)
// and partial sums of exact slot e.
__device__ __forceinline__ Dev prop_view(const Geo& g, const Dev& d, const RepCfg& rc, int s, int e) {
  Dev v = d;
  v.rec = d.recs + s;
  v.mv_idx = d.smv_idx + s * rc.maxm;
  v.mv_old = d.mv_old + s * rc.maxm;
  v.mv_new = d.mv_new + s * rc.maxm;
  v.mv_type = d.mv_type + s * rc.maxm;
  v.tab = d.tab + (size_t)s * rc.maxm * 2 * g.tab_w;
  v.dS = d.dS + (size_t)e * g.n_kh;
  v.part_r = d.part_r + (size_t)e * rc.rstride;
  v.part_mm = d.part_mm + (size_t)e * rc.nmmb;
  return v;
}

// Minimum-image distance of two points in the Ewald box (x, y periodic; z
// periodic in the padded box as the real-space image sum is).
__device__ __forceinline__ double center_dist(const double* a, const double* b, const Geo& g) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double dk = a[k] - b[k];
    dk -= g.L[k] * rint(dk * g.inv_L[k]);
    s += dk * dk;
  }
  return sqrt(s);
}

// ------------------------------------------------------------ proposals
struct PropSmem {
  PrSmem pr;
  double sbuf[kBlock];
  double sh[kBlock / kWave];
  int s_forbid, s_cut, s_prev;
  MoveSel s_sel;
  MoveSel s_sels[kMaxScreen];
};

__global__ __launch_bounds__(kBlock) void k_propose_spec(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                         const McDev* __restrict__ MC,
                                                         const RepCfg* __restrict__ RC, int R) {
  __shared__ PropSmem sm;
  const int p = blockIdx.x, r = blockIdx.y;
  if (r >= R) return;
  const Geo& g = G[r];
  const Dev& d = D[r];
  const McDev& mc = MC[r];
  const RepCfg& rc = RC[r];
  Ctl* C = d.ctl;
  SpecPlan* plan = d.plan;
  const unsigned long long ctr0 = C->move_counter;
  if (threadIdx.x < rc.pmax) sm.s_sels[threadIdx.x] = select_move(d, mc, ctr0 + threadIdx.x);
  __syncthreads();
  if (threadIdx.x == 0) {
    // a molecule proposed again in the batch is proposed from its current
    // coordinates, i.e. on the assumption that its earlier proposal is
    // rejected; the resolve ends the batch there if it was accepted
    int cut = rc.pmax;
    int prev = -1;
    for (int q2 = 0; q2 < p; q2++)
      if (sm.s_sels[q2].mol == sm.s_sels[p].mol) prev = q2;
    sm.s_prev = prev;
    sm.s_sel = sm.s_sels[p];
    if (mc.ctr_limit <= ctr0) cut = 0;
    else if ((long long)(mc.ctr_limit - ctr0) < cut) cut = (int)(mc.ctr_limit - ctr0);
    sm.s_cut = cut;
  }
  __syncthreads();
  const int cut = sm.s_cut;
  if (p == 0 && threadIdx.x == 0) plan->n_prop = cut;
  if (p >= cut) return;
  const int overflow = C->overflow;
  MoveSel sel = sm.s_sel;
  const int was_valid = sel.valid;
  if (overflow) sel.valid = 0;   // the cell list is incomplete: skip, redo later
  const Dev v = prop_view(g, d, rc, p, 0);
  propose_block(g, v, mc, ctr0 + p, sel, sm.pr);
  __syncthreads();
  // spring energy of the trial molecule, wall and dipole parts, bounding
  // sphere: everything of the dE that involves this proposal alone
  const int nm = v.rec->n_moved;
  const int mol = v.rec->mol;
  const int a = d.mol_start[mol], bend = d.mol_start[mol + 1];
  const double bcur = d.bond_cur[mol];
  double ebt = 0;
  // (a valid move that moves no bead keeps the molecule's current energy)
  if (v.rec->valid && g.use_bond && bend - a > 1)
    ebt = molecule_spring_range(g, d, mol, v.rec->lo, v.rec->hi, v.mv_new, sm.sbuf);
  if (threadIdx.x == 0) sm.s_forbid = 0;
  __syncthreads();
  double ew = 0, dmz = 0, cx = 0, cy = 0, cz = 0;
  // bead threadIdx.x kept in registers for the radius pass when every bead
  // has its own thread
  double4 pn0 = make_double4(0, 0, 0, 0), po0 = pn0;
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    const double4 pn = v.mv_new[k], po = v.mv_old[k];
    if (k == threadIdx.x) {
      pn0 = pn;
      po0 = po;
    }
    if (g.use_ext) {
      const WallType w = d.wall[v.mv_type[k]];
      const double en = wall_energy(pn.z, w, g);
      if (en >= kVeryLargeEnergy) atomicOr(&sm.s_forbid, 1);
      ew += en - wall_energy(po.z, w, g);
    }
    dmz += pn.w * (pn.z - po.z);
    cx += pn.x + po.x;
    cy += pn.y + po.y;
    cz += pn.z + po.z;
  }
  // the five sums in one pass (waves, then the wave partials in order, as
  // block_sum does), then the bounding radius (max)
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  double v5[5] = {ew, dmz, cx, cy, cz};
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v5[i] += __shfl_down(v5[i], off, kWave);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 5; i++) sm.sbuf[5 * wv + i] = v5[i];
  __syncthreads();
  double t5[5] = {0, 0, 0, 0, 0};
  for (int w = 0; w < kBlock / kWave; w++)
#pragma unroll
    for (int i = 0; i < 5; i++) t5[i] += sm.sbuf[5 * w + i];
  const double ewt = t5[0], dmzt = t5[1];
  const double inv = nm > 0 ? 1.0 / (2.0 * nm) : 0.0;
  const double c0 = t5[2] * inv, c1 = t5[3] * inv, c2 = t5[4] * inv;
  double rad = 0;
  for (int k = threadIdx.x; k < nm; k += kBlock) {
    const double4 pn = nm <= kBlock ? pn0 : v.mv_new[k], po = nm <= kBlock ? po0 : v.mv_old[k];
    const double dn = (pn.x - c0) * (pn.x - c0) + (pn.y - c1) * (pn.y - c1) + (pn.z - c2) * (pn.z - c2);
    const double dd = (po.x - c0) * (po.x - c0) + (po.y - c1) * (po.y - c1) + (po.z - c2) * (po.z - c2);
    rad = fmax(rad, fmax(dn, dd));
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) rad = fmax(rad, __shfl_down(rad, off, kWave));
  if (lane == 0) sm.sbuf[5 * (kBlock / kWave) + wv] = rad;
  __syncthreads();
  if (threadIdx.x == 0) {
    rad = 0;
    for (int w = 0; w < kBlock / kWave; w++) rad = fmax(rad, sm.sbuf[5 * (kBlock / kWave) + w]);
  }
  if (threadIdx.x == 0) {
    PropInfo& pi = d.prop[p];
    pi.bond_trial = ebt;
    pi.ew = ewt;
    pi.dmz = dmzt;
    pi.c[0] = c0;
    pi.c[1] = c1;
    pi.c[2] = c2;
    pi.rad = sqrt(rad);
    pi.forbid = sm.s_forbid;
    pi.n = nm;
    pi.mol = mol;
    pi.type = v.rec->move_type;
    pi.valid = v.rec->valid;
    pi.skipped = (overflow && was_valid) ? 1 : 0;
    pi.prev = sm.s_prev;
    pi.len = bend - a;
    pi.bcur = bcur;
  }
}

// Plan of a batch, one block per system: first-stage decisions (delayed
// acceptance), exact slots, real-space + moved-moved blocks per exact slot,
// the exact-slot pairs within interaction range (bounding spheres closer than
// the search radius) and their cross-term blocks, the virtual-block count.
constexpr int kPlanThreads = 512;

struct PlanSmem {
  PropInfo s_prop[kMaxScreen];
  double s_nn[kMaxScreen];
  double s_bcur[kMaxScreen];
  double s_u1[kMaxScreen];
  int s_len[kMaxScreen];
  int s_screen[kMaxProp];
  int s_xblk[kMaxPairs];   // cross blocks of a slot pair within range, -1 if not
  int s_pq[kMaxPairs];     // slots (p, q) of pair j: p | q << 8
  int s_ne, s_off;
};
static_assert(kMaxPairs <= 2 * kWave, "near-pair compaction: two pairs per lane of one wave");

// Phase tables of the proposals, in chunks of rows beside the plan (the plan
// is one block per system; the chunks fill the rest of the chip meanwhile).
constexpr int kTabStage = 3072;   // LDS stage of a table chunk, entries

__host__ __device__ __forceinline__ int tab_chunk_rows(int tab_w) {
  const int ws = tab_w | 1;
  return ws <= kTabStage ? min(kTabStage / ws, kPlanThreads / 3) : kPlanThreads / 3;
}

__device__ void plan_tables(const Geo* __restrict__ G, const Dev* __restrict__ D, const RepCfg* __restrict__ RC,
                            int r, int c, int nch, double2* stage) {
  const Geo& g = G[r];
  const Dev& d = D[r];
  const RepCfg& rc = RC[r];
  if (!g.use_ewald || g.n_kh == 0) return;
  const int p = c / nch, ch = c - p * nch;
  if (p >= d.plan->n_prop) return;
  const Dev v = prop_view(g, d, rc, p, 0);
  const int nrow = 2 * v.rec->n_moved, cr = tab_chunk_rows(g.tab_w);
  const int r0 = ch * cr;
  if (r0 >= nrow) return;
  const int ws = (g.tab_w | 1) <= kTabStage ? (g.tab_w | 1) : 0;
  tables_chunk(g, v, r0, min(cr, nrow - r0), stage, ws);
}

// The S(k) of the slots the previous batch accepted, folded in by chunks of
// k vectors beside the plan (before this batch's k-space overwrites their dS
// rows), in slot order.  A batch with no proposal leaves them pending (its
// resolve keeps the mask).
__device__ void plan_fold(const Geo* __restrict__ G, const Dev* __restrict__ D, int r, int f) {
  const Geo& g = G[r];
  const Dev& d = D[r];
  if (d.plan->n_prop == 0) return;
  const unsigned mask = d.plan->acc_mask;
  const int k = f * kPlanThreads + threadIdx.x;
  if (!mask || k >= g.n_kh) return;
  double2 S = d.Sk[k];
  for (int e = 0; e < kMaxProp; e++)
    if (mask & (1u << e)) {
      const double2 ds = d.dS[(size_t)e * g.n_kh + k];
      S.x += ds.x;
      S.y += ds.y;
    }
  d.Sk[k] = S;
}

__global__ __launch_bounds__(kPlanThreads) void k_plan_spec(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                            const McDev* __restrict__ MC,
                            const RepCfg* __restrict__ RC, int* vb_count, int R, int nch, int ntab) {
  __shared__ union {
    PlanSmem p;
    double2 stage[kTabStage];
  } smu;
  PlanSmem& sm = smu.p;
  const int r = blockIdx.x;
  if (r >= R) return;
  if (blockIdx.y > 0) {
    const int c = blockIdx.y - 1;
    if (c < ntab) plan_tables(G, D, RC, r, c, nch, smu.stage);
    else plan_fold(G, D, r, c - ntab);
    return;
  }
  // by-value copies through the constant address space: the per-system
  // structs stay in scalar registers instead of being re-read after stores
  const Geo g = load_const(G, r);
  const Dev d = load_const(D, r);
  const McDev mc = load_const(MC, r);
  const RepCfg rc = load_const(RC, r);
  SpecPlan* plan = d.plan;
  const int np = plan->n_prop;
  const int t = threadIdx.x;
  const unsigned long long ctr0 = d.ctl->move_counter;
  if (t < np) {
    const PropInfo P = d.prop[t];
    sm.s_prop[t] = P;
    sm.s_bcur[t] = P.bcur;
    sm.s_len[t] = P.len;
    const unsigned long long ctr = ctr0 + t;
    const uint4 u = Philox::gen(make_uint4((uint32_t)ctr, (uint32_t)(ctr >> 32), 0xDA1u, 0u),
                                make_uint2((uint32_t)mc.seed, (uint32_t)(mc.seed >> 32)));
    sm.s_u1[t] = u01_32(u.x);
  }
  __syncthreads();
  if (rc.da) {
    // network part of the surrogate, one wave per proposal: sum over moved
    // beads of f(new) - f(old), fixed order
    const int w = t / kWave, lane = t % kWave;
    for (int s = w; s < np; s += kPlanThreads / kWave) {
      const PropInfo& P = sm.s_prop[s];
      const float* f = d.sur_f + (size_t)(2 * s) * rc.maxm;
      double nn = 0;
      if (P.valid)
        for (int m = lane; m < P.n; m += kWave) nn += (double)f[rc.maxm + m] - (double)f[m];
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) nn += __shfl_down(nn, off, kWave);
      if (lane == 0) sm.s_nn[s] = nn;
    }
    __syncthreads();
  }
  if (t < kWave) {
    // one lane per proposal: the first stage, then exact slots in proposal
    // order (a proposal past the last slot is decided next batch), then spare
    // slots for first-stage rejections (training data), by ballots
    const bool have = t < np;
    PropInfo P = {};
    if (have) P = sm.s_prop[t];
    int pass = 1;
    double des = 0;
    if (have && rc.da && P.valid) {
      // surrogate dE: exact spring and wall parts plus the network's estimate
      // of the pair and k-space parts; first stage min(1, exp(-beta dE_s))
      if (g.use_bond && sm.s_len[t] > 1) des += P.bond_trial - sm.s_bcur[t];
      if (g.use_ext) des += P.forbid ? kVeryLargeEnergy : P.ew;
      des += sm.s_nn[t];
      pass = des < kVeryLargeEnergy && sm.s_u1[t] < exp(-g.beta * des);
    }
    const unsigned long long ex = __ballot(have && pass && P.valid);
    const int nex = __popcll(ex);
    int ndec = np;
    if (nex > kMaxProp) {
      // no exact slot left for the (kMaxProp + 1)-th: decided next batch
      unsigned long long m = ex;
      for (int i = 0; i < kMaxProp; i++) m &= m - 1;
      ndec = __builtin_ctzll(m);
    }
    const int ne0 = min(nex, kMaxProp);
    int exact = (have && pass && P.valid && t < ndec) ? lanes_below(ex) : -1;
    const unsigned long long fl = __ballot(have && t < ndec && !pass && P.valid && P.n > 0);
    const int nfill = max(0, min(__popcll(fl), rc.fill - ne0));
    if (((fl >> t) & 1ull) && lanes_below(fl) < nfill) exact = ne0 + lanes_below(fl);
    const int ne = ne0 + nfill;
    if (have) {
      plan->exact_of[t] = exact;
      plan->pass1[t] = pass;
      plan->de_s[t] = des;
      if (exact >= 0) sm.s_screen[exact] = t;
    }
    wave_sync_lds();
    // waves per query for this batch's exact slots (the other systems of a
    // batch taken to have as many), then query and moved-moved blocks per
    // exact slot, prefixes over the slots
    int ncl = t < ne ? 2 * query_clusters(sm.s_prop[sm.s_screen[t]].n) : 0;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) ncl += __shfl_xor(ncl, o, kWave);
    const int qw = query_waves(rc.nsplit, ne * R, ncl, rc.qfill);
    int nq = 0, nmb = 0;
    if (t < ne) {
      const int n = sm.s_prop[sm.s_screen[t]].n;
      nq = n > 0 ? query_blocks(n, qw) : 0;
      nmb = n > 0 ? nmm_blocks(n) : 0;
      plan->screen_of[t] = sm.s_screen[t];
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int u = __shfl_up(nq, o, kWave), v = __shfl_up(nmb, o, kWave);
      if (t >= o) {
        nq += u;
        nmb += v;
      }
    }
    if (t < ne) {
      plan->q_off[t + 1] = nq;
      plan->mm_off[t + 1] = nmb;
    }
    const int off = ne > 0 ? __shfl(nq, ne - 1, kWave) + __shfl(nmb, ne - 1, kWave) : 0;
    if (t == 0) {
      plan->qwave = qw;
      plan->q_off[0] = 0;
      plan->mm_off[0] = 0;
      plan->n_decide = ndec;
      plan->n_exact = ne;
      sm.s_ne = ne;
      sm.s_off = off;
    }
  }
  __syncthreads();
  // slot pairs whose moved beads can interact (bounding spheres within the
  // search radius), one thread per pair
  const int ne = sm.s_ne, npair = ne * (ne - 1) / 2;
  for (int j = t; j < npair; j += kPlanThreads) {
    int e = (int)((1.0 + sqrt(1.0 + 8.0 * j)) * 0.5);
    while (e * (e - 1) / 2 > j) e--;
    while ((e + 1) * e / 2 <= j) e++;
    const int e0 = j - e * (e - 1) / 2;
    const PropInfo& pq = sm.s_prop[sm.s_screen[e]];
    const PropInfo& pp = sm.s_prop[sm.s_screen[e0]];
    int xb = -1;
    if (pq.n > 0 && pp.n > 0 && pp.mol != pq.mol &&
        center_dist(pp.c, pq.c, g) <= pp.rad + pq.rad + g.rs + 1e-9)
      xb = cross_blocks(pp.n, pq.n);
    sm.s_xblk[j] = xb;
    sm.s_pq[j] = e0 | (e << 8);
  }
  __syncthreads();
  if (t >= kWave) return;
  // near pairs in (e, e0) order, e0 < e, compacted by one wave: lane l holds
  // pairs 2l and 2l + 1; a prefix scan gives each its place and block offset
  const int j0 = 2 * t, j1 = j0 + 1;
  const int x0 = j0 < npair ? sm.s_xblk[j0] : -1, x1 = j1 < npair ? sm.s_xblk[j1] : -1;
  const int c0 = x0 >= 0, c1 = x1 >= 0;
  const int own_c = c0 + c1, own_b = (c0 ? x0 : 0) + (c1 ? x1 : 0);
  int cs = own_c, bs = own_b;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int a = __shfl_up(cs, o, kWave), b = __shfl_up(bs, o, kWave);
    if (t >= o) {
      cs += a;
      bs += b;
    }
  }
  int nn = cs - own_c, xo = bs - own_b;
  if (t == 0) plan->pair_off[0] = 0;
  if (c0) {
    plan->pair_p[nn] = sm.s_pq[j0] & 0xff;
    plan->pair_q[nn] = sm.s_pq[j0] >> 8;
    xo += x0;
    nn++;
    plan->pair_off[nn] = xo;
  }
  if (c1) {
    plan->pair_p[nn] = sm.s_pq[j1] & 0xff;
    plan->pair_q[nn] = sm.s_pq[j1] >> 8;
    xo += x1;
    nn++;
    plan->pair_off[nn] = xo;
  }
  if (t == kWave - 1) {
    plan->n_near = cs;
    // k-space blocks: one wave per exact slot and k block
    vb_count[r] = (np > 0 && ne > 0) ? rc.nkb_s * kspace_groups(ne) + sm.s_off + bs : 0;
  }
}

// ------------------------------------------------------------ k space
// One wave per (k block, exact slot), kQWaves slots per block, no block
// barriers.  A k block is kKSpec runs; a run is up to kKRun k vectors of one
// (l_x, l_y) row with consecutive l_z.  Lane (run kk, bead group grp) reads
// the x, y and first z phase and e^{i theta_z} of each of its moved beads at
// both positions from the tables and advances the z phase by multiplication
// (e^{i (l+1) theta} = e^{i l theta} e^{i theta}, conjugated rows included):
// four table reads for up to four k vectors instead of three each.  The bead
// groups of a run meet by wave shuffles; the run's lane writes dS_e(k) and the
// wave sums the slot's dE_e = sum_k c_k (2 Re(S* dS_e) + |dS_e|^2) over the
// block's k vectors (S(k) with the previous batch's acceptances folded in by
// the plan).  The cross sums 2 sum_k c_k Re(dS_p* dS_q) are formed from the
// dS rows by k_reduce_spec.
constexpr int kKRun = 4;
constexpr int kKGrp = kWave / kKSpec;   // bead groups per wave
static_assert(kKSpec * kKGrp == kWave, "a wave is the runs x bead groups of one slot");
static_assert(kMaxProp <= 32, "exact slots are bits of one mask");

__device__ void kspace_spec_wave(const Geo& g, const Dev& d, const RepCfg& rc, const SpecPlan* plan, int kb, int e) {
  const int lane = threadIdx.x % kWave;
  const int kk = lane % kKSpec, grp = lane / kKSpec;
  const int s = uload(&plan->screen_of[e]);
  const int n = uload(&d.recs[s].n_moved);
  const int r = kb * kKSpec + kk;
  const int2 run = r < rc.nrun ? d.krun[r] : make_int2(0, 0);
  const int k0 = run.x, len = run.y;
  TabIdx ti = {0, 0, 0, false, false};
  if (len > 0) ti = tab_index(g, d.kl[k0]);
  const int z1 = g.nkx + 1 + g.nky + 1 + 1;
  double2 acc[kKRun];
#pragma unroll
  for (int i = 0; i < kKRun; i++) acc[i] = make_double2(0, 0);
  if (len > 0) {
    const double4* mo = d.mv_old + s * rc.maxm;
    const double2* base = d.tab + (size_t)s * rc.maxm * 2 * g.tab_w;
    for (int m = grp; m < n; m += kKGrp) {
      const double qm = mo[m].w;
      const double2* to = base + (size_t)(2 * m) * g.tab_w;
      const double2* tn = to + g.tab_w;
      // the eight table reads of both positions first, then new minus old
      double2 ox_n = tn[ti.ox], oy_n = tn[ti.oy], oz_n = tn[ti.oz], b_n = tn[z1];
      double2 ox_o = to[ti.ox], oy_o = to[ti.oy], oz_o = to[ti.oz], b_o = to[z1];
      {
        if (ti.ny) oy_n.y = -oy_n.y;
        if (ti.nz) oz_n.y = -oz_n.y;
        const double2 xy = cmul(ox_n, oy_n);
        double2 z = oz_n;
#pragma unroll
        for (int i = 0; i < kKRun; i++) {
          if (i) z = cmul(z, b_n);
          const double2 ph = cmul(xy, z);
          acc[i].x += qm * ph.x;
          acc[i].y += qm * ph.y;
        }
      }
      {
        if (ti.ny) oy_o.y = -oy_o.y;
        if (ti.nz) oz_o.y = -oz_o.y;
        const double2 xy = cmul(ox_o, oy_o);
        double2 z = oz_o;
#pragma unroll
        for (int i = 0; i < kKRun; i++) {
          if (i) z = cmul(z, b_o);
          const double2 ph = cmul(xy, z);
          acc[i].x -= qm * ph.x;
          acc[i].y -= qm * ph.y;
        }
      }
    }
  }
  // the bead groups of run kk sit in lanes kk, kk + kKSpec, ...; a butterfly
  // leaves their sum in each of them
#pragma unroll
  for (int o = kKSpec; o < kWave; o <<= 1)
#pragma unroll
    for (int i = 0; i < kKRun; i++) {
      acc[i].x += __shfl_xor(acc[i].x, o, kWave);
      acc[i].y += __shfl_xor(acc[i].y, o, kWave);
    }
  double part = 0;
  if (grp == 0) {
#pragma unroll
    for (int i = 0; i < kKRun; i++)
      if (i < len) {
        const double2 v = acc[i], S = d.Sk[k0 + i];
        d.dS[(size_t)e * g.n_kh + k0 + i] = v;
        part += d.ck[k0 + i] * (2.0 * (S.x * v.x + S.y * v.y) + v.x * v.x + v.y * v.y);
      }
  }
  part = wave_sum(part);
  if (lane == 0) d.part_ks[(size_t)e * rc.nkb_s + kb] = part;
}

// ------------------------------------------------------------ real space
// Cluster queries, one per wave (kQWaves per block, no block barriers): a wave
// takes up to kCluster consecutive moved beads of one proposal, all at their
// old or all at their new positions (the query points), and one share of the
// candidate slots of the cells within the search radius of the points'
// bounding box.  The wave lists those cells 64 at a time with a wave prefix
// over their counts; each lane takes candidate slots lane, lane + 64, ...
// (bead index, position and type loaded together, one candidate ahead) and
// tests its candidate against every point, so a candidate's load, cell
// lookup and image decode serve all the points.  The pairs within the radius
// go into a per-wave queue in LDS (ballot + lane prefix); each time the queue
// holds 64 pairs every lane takes one and evaluates its LJ and erfc terms, so
// the lanes never diverge on the pair arithmetic.  A wave reduction sums the
// lanes in fixed order.  Same pair terms as real_query.
constexpr int kQueue = 2 * kWave;   // pair queue of one wave (holds < 2 fills)
static_assert(kCluster <= kWave, "one lane per query point");

template <int P>
struct QWaveSmem {
  int s_cell[kWave];
  int s_img[kWave];
  int s_end[kWave];
  double4 s_pt[kCluster];  // query points (the image nearest the first point) and charges
  int s_ptt[kCluster];     // their types
  double s_r2[kQueue];     // squared distance of a queued pair
  double s_qq[kQueue];     // its charge product
  int s_t[kQueue];         // its LJ pair index | 1 << 24 when LJ applies (no z image in a slab)
  // P = 1: the queue of pairs at or beyond fp64_below, evaluated in FP32
  double s_fr2[P ? kQueue : 1];
  float s_fqq[P ? kQueue : 1];
  int s_ft[P ? kQueue : 1];
};

// LJ + real-space terms of one queued pair (squared distance r2 <= rs^2,
// charge product qq, LJ pair index lji, whether LJ applies), FP64.  Both table
// entries (LJ pair, erfc node) are loaded before either is used, so that a
// queue drain waits for memory once rather than twice in a row: the empty asm
// takes them as inputs, which keeps the loads ahead of both branches.
__device__ __forceinline__ void pair_terms(const Geo& g, const Dev& d, double r2, double qq, int lji, bool lj_ok,
                                           double& elj, double& ere) {
  const double r = sqrt(r2);
  const double x = g.sqrt_alpha * r;
  const int ni = erfc_node_index(x);
  const LJPair lp = d.lj[lji];
  const double2 node = g.erfc_tab[ni];
  asm volatile("" ::"v"(lp.sigma), "v"(lp.cut), "v"(node.x));
  if (lj_ok) elj += lj_energy(r, lp);
  if (g.use_ewald && qq != 0 && r > 0 && r <= g.rc) ere += g.lB * qq * erfc_from_node(x, ni, node) / r;
}

// The same in FP32 (the FP32 pair path, pairs at or beyond fp64_below; the
// cutoff tests compare the FP64 squared distance), added to the FP64 sums.
__device__ __forceinline__ void pair_terms32(const Geo& g, const Dev& d, double r2, float qq, int lji, bool lj_ok,
                                             double& elj, double& ere) {
  const float r2f = (float)r2;
  const float rf = sqrtf(r2f);
  const float x = (float)g.sqrt_alpha * rf;
  const int ni = erfc_node_index32(x);
  const LJPair32 lp = d.lj32[lji];
  const float2 node = g.erfc_tab32[ni];
  asm volatile("" ::"v"(lp.sigma2), "v"(lp.cut), "v"(node.x));
  if (lj_ok) {
    if (r2 < (double)lp.cut * (double)lp.cut) {
      const float s2 = lp.sigma2 / r2f;
      const float r6 = s2 * s2 * s2;
      elj += (double)(lp.eps4 * (r6 * r6 - r6) - lp.shift);
    }
  }
  if (g.use_ewald && qq != 0.f && r2 <= g.rc * g.rc) ere += (double)((float)g.lB * qq * erfc_from_node32(x, ni, node) / rf);
}

// LJ + real-space energy of the np points pts[0 .. np) (charges in .w, types
// ptype[]) against every bead image in the cell list within the search
// radius, skipping the index range [mlo, mhi] (contig) or the moved flags,
// one wave taking share `part` of `nparts` of the candidate slots.  Result
// valid in lane 0.
template <int P>
__device__ double2 real_query_cluster(const Geo& g, const Dev& d, const double4* __restrict__ pts,
                                      const int* __restrict__ ptype, int np, int contig, int mlo, int mhi, int part,
                                      int nparts, QWaveSmem<P>& sm) {
  const int lane = threadIdx.x % kWave;
  // each point as the image nearest the first point's wrapped position (an
  // image of a point has the same pair images within the radius), and the
  // points' bounding box
  double bl[3], bh[3];
  {
    const double4 p0 = pts[0];
    const double p0v[3] = {p0.x, p0.y, p0.z};
    double qv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) qv[k] = wrap_coord(p0v[k], g.L[k], g.inv_L[k]);
    if (lane < np) {
      const double4 p = pts[lane];
      const double pv[3] = {p.x, p.y, p.z};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double dk = pv[k] - p0v[k];
        dk -= g.L[k] * rint(dk * g.inv_L[k]);
        qv[k] += dk;
      }
      sm.s_pt[lane] = make_double4(qv[0], qv[1], qv[2], p.w);
      sm.s_ptt[lane] = ptype[lane];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
      bl[k] = qv[k];
      bh[k] = qv[k];
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        bl[k] = fmin(bl[k], __shfl_xor(bl[k], o, kWave));
        bh[k] = fmax(bh[k], __shfl_xor(bh[k], o, kWave));
      }
    // the butterfly leaves the box in every lane: scalar registers from here
#pragma unroll
    for (int k = 0; k < 3; k++) {
      bl[k] = uniform_double(bl[k]);
      bh[k] = uniform_double(bh[k]);
    }
  }
  int lo[3], nd[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)floor((bl[k] - g.rs) * g.inv_h[k]);
    const int hi = (int)floor((bh[k] + g.rs) * g.inv_h[k]);
    nd[k] = hi - lo[k] + 1;
  }
  const int ncq = nd[0] * nd[1] * nd[2];
  const int nt = g.n_types;
  double elj = 0, ere = 0;
  int qhead = 0, qn = 0;   // queue start and fill (wave-uniform)
  int fhead = 0, fn = 0;   // the same for the FP32 queue (P = 1)
  const double fp64_b2 = P ? g.fp64_below * g.fp64_below : 0.0;
  // the pairs queued at qhead .. qhead + count - 1, one per lane
  auto drain = [&](int count) {
    wave_sync_lds();
    if (lane < count) {
      const int k = (qhead + lane) % kQueue;
      const int t = sm.s_t[k];
      pair_terms(g, d, sm.s_r2[k], sm.s_qq[k], t & 0xffffff, (t >> 24) != 0, elj, ere);
    }
    qhead = (qhead + count) % kQueue;
    qn -= count;
    wave_sync_lds();
  };
  auto drain32 = [&](int count) {
    wave_sync_lds();
    if (lane < count) {
      const int k = (fhead + lane) % kQueue;
      const int t = sm.s_ft[k];
      pair_terms32(g, d, sm.s_fr2[k], sm.s_fqq[k], t & 0xffffff, (t >> 24) != 0, elj, ere);
    }
    fhead = (fhead + count) % kQueue;
    fn -= count;
    wave_sync_lds();
  };
  wave_sync_lds();   // the points are in LDS
  for (int base = 0; base < ncq; base += kWave) {
    const int nloc = min(kWave, ncq - base);
    int end = 0;
    if (lane < nloc) {
      const int c = base + lane;
      const int cz = c % nd[2], cy = (c / nd[2]) % nd[1], cx = c / (nd[2] * nd[1]);
      const int u[3] = {lo[0] + cx, lo[1] + cy, lo[2] + cz};
      int cell = 0, img = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int qk = floordiv(u[k], g.nc[k]);
        const int w = u[k] - qk * g.nc[k];
        cell = cell * g.nc[k] + w;
        img |= ((qk + 128) & 0xff) << (8 * k);
      }
      sm.s_cell[lane] = cell;
      sm.s_img[lane] = img;
      // cells whose box (widened by a rounding margin) lies beyond the search
      // radius of the bounding box hold no candidate
      double dmin2 = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double b0 = (double)u[k] * g.h[k] - 1e-9 * g.L[k], b1 = b0 + g.h[k] + 2e-9 * g.L[k];
        const double dd = fmax(0.0, fmax(b0 - bh[k], bl[k] - b1));
        dmin2 += dd * dd;
      }
      end = dmin2 > g.rs2 ? 0 : d.ccount[cell];
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int v = __shfl_up(end, o, kWave);
      if (lane >= o) end += v;
    }
    sm.s_end[lane] = end;
    const int total = __builtin_amdgcn_readfirstlane(__shfl(end, nloc - 1, kWave));
    wave_sync_lds();
    const int stride = nparts * kWave;
    // candidate it: its cell slot and image (binary search over the wave's
    // cell prefix in LDS)
    auto locate = [&](int it, size_t& si, int& img) {
      int lo_i = 0, hi_i = nloc - 1;
      while (lo_i < hi_i) {
        const int mid = (lo_i + hi_i) >> 1;
        if (sm.s_end[mid] > it) hi_i = mid;
        else lo_i = mid + 1;
      }
      si = (size_t)sm.s_cell[lo_i] * g.cap + (it - (lo_i ? sm.s_end[lo_i - 1] : 0));
      img = sm.s_img[lo_i];
    };
    int it = part * kWave + lane;
    int j = -1, tj = 0, img = 0;
    double4 pj = make_double4(0, 0, 0, 0);
    if (it < total) {
      size_t si;
      locate(it, si, img);
      j = d.cbead[si];
      pj = d.cpos[si];
      tj = d.ctype[si];
    }
    // wave-uniform trip count: the lanes past `total` ride along
    const int rounds = total > part * kWave ? (total - part * kWave + stride - 1) / stride : 0;
    for (int rd = 0; rd < rounds; rd++) {
      const int it2 = it + stride;
      int j2 = -1, tj2 = 0, img2 = 0;
      double4 pj2 = make_double4(0, 0, 0, 0);
      if (it2 < total) {
        size_t si2;
        locate(it2, si2, img2);
        j2 = d.cbead[si2];
        pj2 = d.cpos[si2];
        tj2 = d.ctype[si2];
      }
      // candidate it (its image position) against every point
      const bool ok = it < total && j >= 0 && !(contig ? (j >= mlo && j <= mhi) : d.moved[j]);
      double cx = 0, cy = 0, cz = 0;
      int tflag = 0;
      if (ok) {
        const int iz = ((img >> 16) & 0xff) - 128;
        const int ix = (img & 0xff) - 128, iy = ((img >> 8) & 0xff) - 128;
        cx = pj.x + ix * g.L[0];
        cy = pj.y + iy * g.L[1];
        cz = pj.z + iz * g.L[2];
        tflag = (g.use_pair && (g.npbc == 3 || iz == 0)) ? 1 << 24 : 0;
      }
#pragma unroll 1
      for (int i = 0; i < np; i++) {
        const double4 qp = sm.s_pt[i];
        const double dx = cx - qp.x, dy = cy - qp.y, dz = cz - qp.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const bool in = ok && r2 <= g.rs2;
        // FP64 queue: every pair (P = 0), the close ones (P = 1)
        const bool near = in && (P == 0 || r2 < fp64_b2);
        const unsigned long long m = __ballot(near);
        if (m) {
          if (near) {
            const int k = (qhead + qn + lanes_below(m)) % kQueue;
            sm.s_r2[k] = r2;
            sm.s_qq[k] = qp.w * pj.w;
            sm.s_t[k] = (sm.s_ptt[i] * nt + tj) | tflag;
          }
          qn += __popcll(m);
          if (qn >= kWave) drain(kWave);
        }
        if (P == 1) {
          // each queue drains 64 pairs of one precision at a time, so the
          // lanes of a drain never split between the two
          const bool far = in && !near;
          const unsigned long long mf = __ballot(far);
          if (mf) {
            if (far) {
              const int k = (fhead + fn + lanes_below(mf)) % kQueue;
              sm.s_fr2[k] = r2;
              sm.s_fqq[k] = (float)(qp.w * pj.w);
              sm.s_ft[k] = (sm.s_ptt[i] * nt + tj) | tflag;
            }
            fn += __popcll(mf);
            if (fn >= kWave) drain32(kWave);
          }
        }
      }
      it = it2;
      j = j2;
      tj = tj2;
      img = img2;
      pj = pj2;
    }
    wave_sync_lds();   // the next chunk rewrites the cell arrays
  }
  if (qn > 0) drain(qn);
  if (P == 1 && fn > 0) drain32(fn);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    elj += __shfl_down(elj, off, kWave);
    ere += __shfl_down(ere, off, kWave);
  }
  return make_double2(elj, ere);
}

// Cross block: LJ + real pairs between slot q's and slot p's moved beads,
// u(i',j') - u(i,j') - u(i',j) + u(i,j): the change of either one's dE when
// the other has been accepted (same pair functions as the moved-moved pairs).
__device__ void cross_block(const Geo& g, const Dev& d, const RepCfg& rc, int sp, int sq, int np_, int nq,
                            int local, double* sh, double2* out) {
  const int t = local * kBlock + threadIdx.x;
  double elj = 0, ere = 0;
  if (t < np_ * nq) {
    const int i = t / np_, j = t - (t / np_) * np_;
    const int si = sq * rc.maxm + i, sj = sp * rc.maxm + j;
    const double4 a0 = d.mv_old[si], a1 = d.mv_new[si], b0 = d.mv_old[sj], b1 = d.mv_new[sj];
    const double ao[3] = {a0.x, a0.y, a0.z}, an[3] = {a1.x, a1.y, a1.z};
    const double bo[3] = {b0.x, b0.y, b0.z}, bn[3] = {b1.x, b1.y, b1.z};
    if (g.use_pair) {
      const LJPair lp = d.lj[d.mv_type[si] * g.n_types + d.mv_type[sj]];
      elj = lj_energy(bb_dist(an, bn, g.Lsim, g.npbc), lp) - lj_energy(bb_dist(ao, bn, g.Lsim, g.npbc), lp) -
            lj_energy(bb_dist(an, bo, g.Lsim, g.npbc), lp) + lj_energy(bb_dist(ao, bo, g.Lsim, g.npbc), lp);
    }
    if (g.use_ewald && a0.w * b0.w != 0)
      ere = real_pair(an, a1.w, bn, b1.w, g) - real_pair(ao, a0.w, bn, b1.w, g) -
            real_pair(an, a1.w, bo, b0.w, g) + real_pair(ao, a0.w, bo, b0.w, g);
  }
  const double a = block_sum<kBlock>(elj, sh);
  const double c = block_sum<kBlock>(ere, sh);
  if (threadIdx.x == 0) *out = make_double2(a, c);
}

struct CSpecSmem {
  double sh[kBlock / kWave];
};

template <int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_compute_spec(
    const Geo* __restrict__ G, const Dev* __restrict__ D, const RepCfg* __restrict__ RC,
    const int* __restrict__ vb_count, int R) {
  __shared__ int s_pre[kMaxReplicas];
  __shared__ struct U {
    QWaveSmem<P> q[kQWaves];
  } sm;
  __shared__ CSpecSmem cs;
  for (int r = threadIdx.x; r < R; r += kBlock) s_pre[r] = vb_count[r];
  __syncthreads();
  lds_scan(s_pre, R);
  const int total = s_pre[R - 1];
  for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const int r = __builtin_amdgcn_readfirstlane(find_replica(s_pre, R, vb));
    int b = __builtin_amdgcn_readfirstlane(vb - (r ? s_pre[r - 1] : 0));
    const Geo g = load_const(G, r);
    const Dev d = load_const(D, r);
    const RepCfg rc = load_const(RC, r);
    const SpecPlan* plan = d.plan;
    const int ne = uload(&plan->n_exact);
    const int ngrp = kspace_groups(ne);
    // longest work first: queries, k space, moved-moved pairs, cross terms
    const int nqt = uload(&plan->q_off[ne]);
    if (b < nqt) {
      int e = 0;
      while (e + 1 < ne && uload(&plan->q_off[e + 1]) <= b) e++;
      const int local = b - uload(&plan->q_off[e]);
      const Dev v = prop_view(g, d, rc, uload(&plan->screen_of[e]), e);
      const int nm = uload(&v.rec->n_moved);
      const int qw = uload(&plan->qwave);
      // wave gw: share gw % qw of query q2 = gw / qw = 2 c + on, cluster c
      // (moved beads c kCluster ..) at their old (on = 0) or new positions
      const int w = threadIdx.x / kWave, gw = local * kQWaves + w;
      if (gw < 2 * query_clusters(nm) * qw) {
        const int q2 = gw / qw, part = gw - q2 * qw;
        const int m0 = (q2 >> 1) * kCluster, on = q2 & 1;
        const double2 eq = real_query_cluster<P>(g, v, (on ? v.mv_new : v.mv_old) + m0, v.mv_type + m0,
                                                 min(kCluster, nm - m0), v.rec->contig, v.rec->lo, v.rec->hi, part,
                                                 qw, sm.q[w]);
        const double sgn = on ? 1.0 : -1.0;
        if (threadIdx.x % kWave == 0) v.part_r[gw] = make_double2(sgn * eq.x, sgn * eq.y);
      }
    } else if ((b -= nqt) < rc.nkb_s * ngrp) {
      // k block b / ngrp, exact slots kQWaves (b % ngrp) + wave
      const int kb = b / ngrp, e = (b - kb * ngrp) * kQWaves + threadIdx.x / kWave;
      if (e < ne) kspace_spec_wave(g, d, rc, plan, kb, e);
    } else {
      b -= rc.nkb_s * ngrp;
      const int nmt = uload(&plan->mm_off[ne]);
      if (b < nmt) {
        int e = 0;
        while (e + 1 < ne && uload(&plan->mm_off[e + 1]) <= b) e++;
        const Dev v = prop_view(g, d, rc, uload(&plan->screen_of[e]), e);
        const int nm = uload(&v.rec->n_moved);
        mm_block(g, v, b - uload(&plan->mm_off[e]), nmm_blocks(nm), cs.sh);
      } else {
        b -= nmt;
        const int nn = uload(&plan->n_near);
        int j = 0;
        while (j + 1 < nn && uload(&plan->pair_off[j + 1]) <= b) j++;
        const int off = uload(&plan->pair_off[j]);
        const int sp = uload(&plan->screen_of[uload(&plan->pair_p[j])]);
        const int sq = uload(&plan->screen_of[uload(&plan->pair_q[j])]);
        const int np_ = uload(&d.recs[sp].n_moved), nq = uload(&d.recs[sq].n_moved);
        cross_block(g, d, rc, sp, sq, np_, nq, b - off, cs.sh, d.part_x + b);
      }
    }
    __syncthreads();  // the LDS is reused by the next virtual block
  }
}

constexpr int kReduceThreads = 1024;

// Sums of a batch, one block per (sum, system), fixed order: blockIdx.x in
// [0, kMaxProp) the k-space dE per exact slot (the k blocks' partials), then
// kMaxPairs k-space cross sums 2 sum_k c_k Re(dS_p* dS_q) of the slot pairs
// p < q (pair_index order) from the dS rows, then kMaxProp (LJ, real) sums per
// exact slot (real-space query blocks + moved-moved blocks), then kMaxPairs
// real-space cross sums of the near pairs.
__global__ __launch_bounds__(kReduceThreads) void k_reduce_spec(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                        const RepCfg* __restrict__ RC, int R) {
  __shared__ double sh[kReduceThreads / kWave];
  const int j = blockIdx.x, r = blockIdx.y;
  if (r >= R) return;
  const Geo& g = G[r];
  const Dev& d = D[r];
  const RepCfg& rc = RC[r];
  SpecPlan* plan = d.plan;
  if (plan->n_prop == 0) return;
  const int ne = plan->n_exact;
  const int t = threadIdx.x;
  if (j < kMaxProp) {
    if (j >= ne) return;
    double s = 0;
    for (int i = t; i < rc.nkb_s; i += kReduceThreads) s += d.part_ks[(size_t)j * rc.nkb_s + i];
    s = block_sum<kReduceThreads>(s, sh);
    if (t == 0) plan->ksum[j] = s;
  } else if (j < kKVals) {
    const int pj = j - kMaxProp;
    int q = (int)((1.0 + sqrt(1.0 + 8.0 * pj)) * 0.5);
    while (q * (q - 1) / 2 > pj) q--;
    while ((q + 1) * q / 2 <= pj) q++;
    const int p = pj - q * (q - 1) / 2;
    if (q >= ne) return;
    double s = 0;
    if (rc.nkb_s > 0) {
      const double2* a = d.dS + (size_t)p * g.n_kh;
      const double2* b = d.dS + (size_t)q * g.n_kh;
#pragma unroll 4
      for (int k = t; k < g.n_kh; k += kReduceThreads) {
        const double2 u = a[k], v = b[k];
        s += d.ck[k] * (u.x * v.x + u.y * v.y);
      }
    }
    s = block_sum<kReduceThreads>(s, sh);
    if (t == 0) plan->ksum[j] = 2.0 * s;
  } else if (j < kKVals + kMaxProp) {
    const int e = j - kKVals;
    if (e >= ne) return;
    const int nm = d.recs[plan->screen_of[e]].n_moved;
    if (nm == 0) return;
    const int nreal = 2 * query_clusters(nm) * plan->qwave, nmm = nmm_blocks(nm);
    const double2* pr = d.part_r + (size_t)e * rc.rstride;
    const double2* pm = d.part_mm + (size_t)e * rc.nmmb;
    double lj = 0, re = 0;
    for (int i = t; i < nreal; i += kReduceThreads) {
      lj += pr[i].x;
      re += pr[i].y;
    }
    for (int i = t; i < nmm; i += kReduceThreads) {
      lj += pm[i].x;
      re += pm[i].y;
    }
    lj = block_sum<kReduceThreads>(lj, sh);
    re = block_sum<kReduceThreads>(re, sh);
    if (t == 0) plan->rsum[e] = make_double2(lj, re);
  } else {
    const int jj = j - kKVals - kMaxProp;
    if (jj >= plan->n_near) return;
    const int off = plan->pair_off[jj], nblk = plan->pair_off[jj + 1] - off;
    double lj = 0, re = 0;
    for (int i = t; i < nblk; i += kReduceThreads) {
      lj += d.part_x[off + i].x;
      re += d.part_x[off + i].y;
    }
    lj = block_sum<kReduceThreads>(lj, sh);
    re = block_sum<kReduceThreads>(re, sh);
    if (t == 0) plan->xsum[jj] = make_double2(lj, re);
  }
}

// ------------------------------------------------------------ resolve
// The items of consecutive accepted exact slots committed together:
// group slot j holds items [goff[j], goff[j + 1]) of proposal gscr[j].
struct SlotItems {
  const double4* mv_new;
  const int* mv_idx;
  int maxm;
  const int* goff;
  const int* gscr;
  int ns, n;
  __device__ __forceinline__ void get(int k, double4& p, int& i) const {
    int j = 0;
    while (j + 1 < ns && goff[j + 1] <= k) j++;
    const size_t o = (size_t)gscr[j] * maxm + (k - goff[j]);
    p = mv_new[o];
    i = mv_idx[o];
  }
};

struct ResSmem {
  CommitSmem cm;
  PropInfo s_prop[kMaxScreen];
  double s_bcur[kMaxScreen];
  int s_len[kMaxScreen];
  double s_u[kMaxScreen];
  double s_des[kMaxScreen];
  int s_pass[kMaxScreen];
  int s_exact[kMaxScreen];
  int s_screen[kMaxProp];
  double2 s_rsum[kMaxProp];
  double s_ksum[kKVals];
  double2 s_xsum[kMaxPairs];
  int s_near[kMaxPairs];
  int s_train[kMaxProp];         // ring slot of a training sample per exact slot, -1 none
  int s_drec[kMaxScreen];        // decision-record index of a proposal, -1 none
  int s_goff[kMaxProp + 1], s_gscr[kMaxProp];
  int s_mask, s_ndone;
  double s_Mz, s_dipl, s_run[8];   // control block: dipole moment, dipole lag, running totals
};

static_assert(kMaxScreen <= kWave, "one lane per proposal in the decision wave");

__device__ __forceinline__ double lane_double(double v, int lane) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The reference's sequential Metropolis chain over the batch
// (simulation.cc:241-355), one wave, one lane per proposal.  A pass
// evaluates every undecided proposal against the acceptances so far; the
// proposals before the first acceptance (or stop) of the pass are decided
// as they would be one after the other, since nothing they read changes
// until then.  The accepted one updates the cross terms, the dipole moment
// and the stop conditions of the proposals after it, and the next pass
// starts behind it: one pass per acceptance instead of one step per
// proposal.  Each dE adds the cross terms in acceptance order, which is the
// order of increasing exact slot, as the one-at-a-time chain does.
__device__ void resolve_decide(const Geo& g, const Dev& d, const RepCfg& rc, Ctl* C, SpecPlan* plan, int np,
                               int ndec, unsigned long long ctr0, ResSmem& sm) {
  const int lane = threadIdx.x;
  const bool mine = lane < np;
  PropInfo P = {};
  P.prev = -1;
  P.type = -1;
  if (mine) P = sm.s_prop[lane];
  const int e = mine ? sm.s_exact[lane] : -1;
  const int pass = mine ? sm.s_pass[lane] : 0;
  const double des = mine ? sm.s_des[lane] : 0.0, u = mine ? sm.s_u[lane] : 1.0;
  const int len = mine ? sm.s_len[lane] : 0;
  const double bcur = mine ? sm.s_bcur[lane] : 0.0;
  const int use_pair = g.use_pair, use_ext = g.use_ext, use_ewald = g.use_ewald, use_bond = g.use_bond;
  const int dipole = g.dipole, dipole_exact = g.dipole_exact, da_on = rc.da, train = rc.train;
  const double dip_coef = g.dip_coef, beta = g.beta;
  // pair and k-space dE against the state before the batch (a proposal that
  // moves no bead has no real-space blocks: its pair terms are zero), plus
  // the cross terms of the acceptances so far
  double lj = 0, re = 0, kk = 0;
  if (e >= 0) {
    lj = P.n > 0 ? sm.s_rsum[e].x : 0.0;
    re = P.n > 0 ? sm.s_rsum[e].y : 0.0;
    kk = sm.s_ksum[e];
  }
  const bool evaluated = P.valid && (pass || e >= 0);
  bool stop_prev = false, stale = false;
  unsigned mask = 0;
  double Mz = sm.s_Mz, cur_dipl = sm.s_dipl;
  double run[8];
  for (int k = 0; k < 8; k++) run[k] = sm.s_run[k];
  // what each proposal was decided with
  double f_tt[8] = {0, 0, 0, 0, 0, 0, 0, 0}, f_dE = 0, f_target = 0;
  bool f_acc = false, f_early = false;
  int done = ndec, cut = 0, s = 0;
  while (s < ndec) {
    // ForceField::EnergyDifference with its early exits (force_field.cc:407-434)
    double tt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double dE = 0;
    bool early = false;
    if (use_pair) {
      tt[T_LJ] = lj;
      dE += lj;
      if (dE >= kVeryLargeEnergy) early = true;
    }
    if (!early && use_ext) {
      tt[T_WALL] = P.forbid ? kVeryLargeEnergy : P.ew;
      dE += tt[T_WALL];
      if (dE >= kVeryLargeEnergy) early = true;
    }
    double trial_dipl = cur_dipl;
    if (!early && use_ewald) {
      tt[T_REAL] = re;
      tt[T_KSPACE] = kk;
      double dd = 0;
      if (dipole) {
        if (dipole_exact) {
          const double mz1 = Mz + P.dmz;
          trial_dipl = dipole_energy(dip_coef, mz1);
          dd = dipole_change(trial_dipl, dipole_energy(dip_coef, Mz));
        } else {
          // potential_ewald.cc:528-531: dipole of the CURRENT coordinates
          trial_dipl = dipole_energy(dip_coef, Mz);
          dd = dipole_change(trial_dipl, cur_dipl);
        }
      }
      tt[T_DIPOLE] = dd;
      dE += re + kk + dd;
    }
    if (!early && use_bond && len > 1) {
      tt[T_BOND] = P.bond_trial - bcur;
      dE += tt[T_BOND];
    }
    // Metropolis (simulation.cc:327-332), no draw for a forbidden move; with
    // the surrogate, the second stage min(1, exp(-beta (dE - dE_s)))
    const bool acc = evaluated && pass && (dE < kVeryLargeEnergy) && (u < exp(-beta * (dE - des)));
    const bool stop = stop_prev || stale;
    const bool open = lane >= s && lane < ndec;
    const unsigned long long ev = __ballot(open && (stop || acc));
    const int f = ev ? __builtin_ctzll(ev) : ndec;
    if (open && (lane < f || (lane == f && !stop))) {
      for (int k = 0; k < 8; k++) f_tt[k] = tt[k];
      f_dE = dE;
      f_early = early;
      f_acc = lane == f;
      f_target = lj + re + kk;
    }
    if (f >= ndec) break;
    if (__builtin_amdgcn_readlane((int)stop, f)) {
      // a proposal made on the assumption that its molecule's earlier
      // proposal in the batch is rejected, or whose surrogate value could see
      // an acceptance of this batch: it is made again next batch
      done = f;
      cut = __builtin_amdgcn_readlane((int)(stale && !stop_prev), f);
      break;
    }
    // proposal f is accepted
    const int ef = __builtin_amdgcn_readlane(e, f);
    mask |= 1u << ef;
    Mz += lane_double(P.dmz, f);
    if (dipole) cur_dipl = lane_double(trial_dipl, f);
    for (int k = 0; k < 8; k++) run[k] += lane_double(tt[k], f);
    if (lane == f && use_bond && len > 1) d.bond_cur[P.mol] = P.bond_trial;
    const double fc0 = lane_double(P.c[0], f), fc1 = lane_double(P.c[1], f), fc2 = lane_double(P.c[2], f);
    const double frad = lane_double(P.rad, f);
    if (lane > f && lane < ndec) {
      if (e >= 0) {
        const int pi = pair_index(ef, e);
        kk += sm.s_ksum[kMaxProp + pi];
        const int j = sm.s_near[pi];
        if (j >= 0) {
          lj += sm.s_xsum[j].x;
          re += sm.s_xsum[j].y;
        }
      }
      if (P.prev == f) stop_prev = true;
      if (da_on) {
        const double fc[3] = {fc0, fc1, fc2};
        stale = stale || center_dist(fc, P.c, g) <= frad + P.rad + (double)rc.r_d + 1e-9;
      }
    }
    s = f + 1;
  }
  // counters of the decided proposals
  const bool decided = lane < done;
  const bool valid = decided && P.valid;
  long long att[5], acc5[5];
  for (int k = 0; k < 5; k++) {
    att[k] = __popcll(__ballot(valid && P.type == k));
    acc5[k] = __popcll(__ballot(valid && f_acc && P.type == k));
  }
  const long long n_valid = __popcll(__ballot(valid));
  const long long skipped = __popcll(__ballot(decided && !P.valid && P.skipped));
  const long long s1r = __popcll(__ballot(valid && !pass));
  const long long tev = __popcll(__ballot(valid && !pass && e >= 0));
  const long long s1p = da_on ? __popcll(__ballot(valid && pass)) : 0;
  if (d.drec) {
    // bounded record of the decided proposals, in decision order: the terms
    // and total the Metropolis test was given (trial positions copied by
    // the block afterwards)
    const unsigned long long rm = __ballot(valid);
    const int rhead = d.drec_ctl[0], rcap = d.drec_ctl[1];
    int slot = -1;
    if (valid) {
      const int k = rhead + lanes_below(rm);
      if (k < rcap) {
        slot = k;
        DecRec& o = d.drec[k];
        o.move = ctr0 + lane;
        o.mol = P.mol;
        o.n = P.n;
        o.type = P.type;
        o.accepted = f_acc ? 1 : 0;
        o.lo = d.recs[lane].lo;
        o.hi = d.recs[lane].hi;
        o.contig = d.recs[lane].contig;
        o.evaluated = evaluated ? 1 : 0;
        o.early = f_early ? 1 : 0;
        o.pass = pass;
        o.pad0 = o.pad1 = 0;
        for (int k2 = 0; k2 < 8; k2++) o.terms[k2] = f_tt[k2];
        o.total = f_dE;
        o.de_s = des;
      }
    }
    if (mine) sm.s_drec[lane] = slot;
    if (lane == 0) d.drec_ctl[0] = rhead + (int)__popcll(rm);
  }
  // training samples in proposal order: the pair + k-space part the network
  // estimates
  int head = train ? d.tr_ctl[0] : 0;
  const int ring = train ? d.tr_ctl[1] : 0;
  const bool sample = valid && e >= 0 && train && !f_early && P.n > 0 && ring > 0;
  const unsigned long long smask = __ballot(sample);
  if (sample) {
    const int slot = (head + lanes_below(smask)) % ring;
    d.tr_target[slot] = f_target;
    d.tr_n[slot] = P.n;
    sm.s_train[e] = slot;
  }
  head += __popcll(smask);
  // the terms of the last proposal that reached the Metropolis test
  const unsigned long long lm = __ballot(valid && pass);
  double last[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double last_total = C->total;
  if (lm) {
    const int l = 63 - __builtin_clzll(lm);
    for (int k = 0; k < 8; k++) last[k] = lane_double(f_tt[k], l);
    last_total = lane_double(f_dE, l);
  }
  if (lane == 0) {
    C->Mz = Mz;
    C->cur_dipl = cur_dipl;
    C->trial_dipl = cur_dipl;
    for (int k = 0; k < 8; k++) {
      C->running[k] = run[k];
      C->terms[k] = last[k];
    }
    C->total = last_total;
    for (int k = 0; k < 5; k++) {
      C->attempted[k] += att[k];
      C->accepted[k] += acc5[k];
    }
    C->skipped += skipped;
    C->valid_moves += n_valid;
    C->s1_rejected += s1r;
    C->s1_passed += s1p;
    C->train_evals += tev;
    C->stale_cuts += cut;
    C->move_counter = ctr0 + done;
    if (train) d.tr_ctl[0] = head;
    plan->acc_mask = mask;
    sm.s_mask = mask;
    sm.s_ndone = done;
  }
}

__global__ __launch_bounds__(kBlock) void k_resolve_spec(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                         const McDev* __restrict__ MC,
                                                         const RepCfg* __restrict__ RC, int R) {
  __shared__ ResSmem sm;
  const int r = blockIdx.x;
  if (r >= R) return;
  const Geo& g = G[r];
  const Dev& d = D[r];
  const McDev& mc = MC[r];
  const RepCfg& rc = RC[r];
  Ctl* C = d.ctl;
  SpecPlan* plan = d.plan;
  const int np = plan->n_prop;
  if (np == 0) return;
  const int ndec = plan->n_decide;
  const unsigned long long ctr0 = C->move_counter;
  // everything the decisions read, fetched in parallel
  const int t = threadIdx.x;
  if (t < np) {
    const PropInfo P = d.prop[t];
    sm.s_prop[t] = P;
    sm.s_bcur[t] = P.bcur;
    sm.s_len[t] = P.len;
    sm.s_des[t] = plan->de_s[t];
    sm.s_pass[t] = plan->pass1[t];
    sm.s_exact[t] = plan->exact_of[t];
    const unsigned long long ctr = ctr0 + t;
    const uint4 u = Philox::gen(make_uint4((uint32_t)ctr, (uint32_t)(ctr >> 32), 0xACCu, 0u),
                                make_uint2((uint32_t)mc.seed, (uint32_t)(mc.seed >> 32)));
    sm.s_u[t] = u01_32(u.x);
  }
  if (t >= 64 && t < 64 + kKVals) sm.s_ksum[t - 64] = plan->ksum[t - 64];
  if (t >= 128 && t < 128 + kMaxPairs) {
    const int j = t - 128;
    sm.s_near[j] = -1;
    sm.s_xsum[j] = plan->xsum[j];
  }
  if (t >= 192 && t < 192 + kMaxProp) {
    sm.s_rsum[t - 192] = plan->rsum[t - 192];
    sm.s_screen[t - 192] = plan->screen_of[t - 192];
    sm.s_train[t - 192] = -1;
  }
  if (t >= 224 && t < 224 + kMaxScreen) sm.s_drec[t - 224] = -1;
  if (t == kBlock - 1) {
    sm.s_Mz = C->Mz;
    sm.s_dipl = C->cur_dipl;
  }
  if (t >= kBlock - 9 && t < kBlock - 1) sm.s_run[t - (kBlock - 9)] = C->running[t - (kBlock - 9)];
  __syncthreads();
  {
    const int nn = plan->n_near;
    for (int j = t; j < nn; j += kBlock) sm.s_near[pair_index(plan->pair_p[j], plan->pair_q[j])] = j;
  }
  __syncthreads();
  if (t < kWave) resolve_decide(g, d, rc, C, plan, np, ndec, ctr0, sm);
  __syncthreads();
  if (d.drec) {
    // trial positions and bead indices of the recorded proposals
    for (int s = 0; s < np; s++) {
      const int slot = sm.s_drec[s];
      if (slot < 0) continue;
      const int n = sm.s_prop[s].n;
      for (int i = t; i < n; i += kBlock) {
        d.drec_pos[(size_t)slot * rc.maxm + i] = d.mv_new[(size_t)s * rc.maxm + i];
        d.drec_idx[(size_t)slot * rc.maxm + i] = d.smv_idx[(size_t)s * rc.maxm + i];
      }
    }
  }
  // cell-list commits of the accepted slots, as many slots at a time as the
  // item arrays hold
  const unsigned mask = sm.s_mask;
  int e0 = 0;
  while (e0 < kMaxProp) {
    if (e0 > 0) __syncthreads();   // the last group's commit is done with the group arrays
    int ns = 0, n = 0, e = e0;
    for (; e < kMaxProp; e++) {
      if (!(mask & (1u << e))) continue;
      const int nm = sm.s_prop[sm.s_screen[e]].n;
      if (ns > 0 && n + nm > kMaxChainDev) break;
      if (t == 0) {
        sm.s_goff[ns] = n;
        sm.s_gscr[ns] = sm.s_screen[e];
      }
      ns++;
      n += nm;
    }
    e0 = e;
    if (ns == 0) break;
    if (t == 0) sm.s_goff[ns] = n;
    __syncthreads();
    commit_cells(g, d, SlotItems{d.mv_new, d.smv_idx, rc.maxm, sm.s_goff, sm.s_gscr, ns, n}, sm.cm);
  }
  if (rc.train) {
    // descriptor rows of the training samples (old rows, then new rows)
    for (int e = 0; e < kMaxProp; e++) {
      const int slot = sm.s_train[e];
      if (slot < 0) continue;
      const int s = sm.s_screen[e];
      const int n = sm.s_prop[s].n;
      const float* src = d.sur_x + (size_t)(2 * s) * rc.maxm * kDescr;
      float* dst = d.tr_x + (size_t)slot * 2 * rc.maxm * kDescr;
      for (int i = threadIdx.x; i < n * kDescr; i += kBlock) {
        dst[i] = src[i];
        dst[(size_t)rc.maxm * kDescr + i] = src[(size_t)rc.maxm * kDescr + i];
      }
    }
  }
}

// Folds the pending accepted dS_e(k) into S(k) (end of a run); one block row
// per system, grid-stride over its k vectors.
__global__ void k_flush_spec(const Geo* __restrict__ G, const Dev* __restrict__ D, int R) {
  const int r = blockIdx.y;
  if (r >= R) return;
  const Geo& g = G[r];
  const Dev& d = D[r];
  const unsigned mask = d.plan->acc_mask;
  if (!mask) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < g.n_kh; k += gridDim.x * blockDim.x) {
    double2 s = d.Sk[k];
    for (int e = 0; e < kMaxProp; e++)
      if (mask & (1u << e)) {
        const double2 ds = d.dS[(size_t)e * g.n_kh + k];
        s.x += ds.x;
        s.y += ds.y;
      }
    d.Sk[k] = s;
  }
}

__global__ void k_clear_mask(const Dev* __restrict__ D, int R) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) D[r].plan->acc_mask = 0;
}

// Tombstone compaction of every system's cells; cell_off is the exclusive
// prefix of the cell counts (R + 1 entries).
__global__ void k_compact_batch(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                const int* __restrict__ cell_off, int R) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= cell_off[R]) return;
  int lo = 0, hi = R - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cell_off[mid] <= gid) lo = mid;
    else hi = mid - 1;
  }
  compact_cell(G[lo], D[lo], gid - cell_off[lo]);
}

// Per system: overflow flag and moves still to propose (read by the host
// between graph replays).
__global__ void k_batch_status(const Dev* __restrict__ D, const McDev* __restrict__ MC, int R, int2* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const unsigned long long c = D[r].ctl->move_counter, lim = MC[r].ctr_limit;
  out[r] = make_int2(D[r].ctl->overflow, c < lim ? (int)min(lim - c, 1ull << 30) : 0);
}

}  // namespace plum
