// Surrogate of the trial-move dE for delayed acceptance: per moved bead a
// local descriptor (radial basis functions of the distances to the unmoved
// beads within r_d, one set per sign of q_i q_j, plus the bead's charge and
// its distances to the walls), a 64-128-128-128-1 tanh network f, and
//   dE_nn = sum over moved beads of f(descriptor at new) - f(descriptor at old)
// (antisymmetric in old/new by construction).  Inference runs in BF16 on the
// matrix cores (v_mfma_f32_32x32x8bf16_1k, FP32 accumulation); training runs
// in FP32 (forward, backward and Adam kernels below).  No reference
// counterpart: the reference evaluates every trial move exactly.
#pragma once
#include "kernels_batch.h"

namespace plum {

constexpr int kRbf = 16;                    // radial basis functions per channel
constexpr int kChan = 3;                    // q_i q_j > 0, < 0, = 0
constexpr int kHid = 128;                   // hidden width
constexpr float kRbfMin = 0.7f;             // first basis centre (sigma)
constexpr float kFeatScale = 0.25f;         // feature scale
// parameter layout (FP32 master copy, row-major, W[in][out])
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kDescr * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kHid;
constexpr int kNParams = kOffB4 + 1;

// Smooth cutoff and basis: phi_k(r) = exp(-gamma (r - mu_k)^2) (cos(pi r / r_d) + 1) / 2.
__host__ __device__ __forceinline__ float rbf_gamma(float r_d) {
  const float dmu = (r_d - kRbfMin) / (kRbf - 1);
  return 1.0f / (dmu * dmu);
}

// ------------------------------------------------------------ descriptors
constexpr int kDescrWaves = kBlock / kWave;
constexpr int kDescrLoads = 4;                       // candidate loads in flight per lane
constexpr int kDescrQueue = kDescrLoads * kWave;     // in-range candidates of one round

// Per-wave scratch of descr_wave: one chunk of up to 64 query cells, and the
// in-range candidates of one round of the scan.
struct DescrWaveSmem {
  int s_cell[kWave];
  int s_img[kWave];
  int s_end[kWave];
  float s_rel[kWave][3];
  float2 s_q[kDescrQueue];   // (r, smooth cutoff) of an in-range candidate
  int s_qc[kDescrQueue];     // its channel
};

// Descriptor sums of moved bead m of proposal view v at its old (on = 0) or
// new (on = 1) position, against every unmoved bead image within r_d (images
// as the LJ term takes them: z unshifted with two periodic dimensions), over
// the candidates of wave wv of nw that share the row.  No block barriers.
// The wave scans the query cells 64 at a time, kDescrLoads candidates per
// lane per round; the candidates within r_d (a small fraction of those
// scanned) go to a queue, and lane f then adds basis function f of every
// queued candidate of its channel, in queue order: the sum of feature f ends
// in lane f with one exponential per candidate and lane, and no lane runs
// the basis of a candidate it does not need.
__device__ float descr_wave(const Geo& g, const Dev& v, float r_d, const double4& p, int wv, int nw,
                            DescrWaveSmem& sm) {
  const int lane = threadIdx.x % kWave;
  const int contig = v.rec->contig, mlo = v.rec->lo, mhi = v.rec->hi;
  const double pw[3] = {wrap_coord(p.x, g.L[0], g.inv_L[0]), wrap_coord(p.y, g.L[1], g.inv_L[1]),
                        wrap_coord(p.z, g.L[2], g.inv_L[2])};
  int lo[3], nd[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = (int)floor((pw[k] - r_d) * g.inv_h[k]);
    nd[k] = (int)floor((pw[k] + r_d) * g.inv_h[k]) - lo[k] + 1;
  }
  const int ncq = nd[0] * nd[1] * nd[2];
  const float qi = (float)p.w;
  const float rd2 = r_d * r_d, gam = rbf_gamma(r_d), dmu = (r_d - kRbfMin) / (kRbf - 1);
  const float pi_rd = 3.14159265358979f / r_d;
  // this lane's feature: channel and basis centre (lanes past the basis
  // features add nothing)
  const int my_ch = lane < kRbf * kChan ? lane / kRbf : -1;
  const float my_mu = kRbfMin + (lane % kRbf) * dmu;
  float acc = 0.f;
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
        const int q = floordiv(u[k], g.nc[k]);
        const int w = u[k] - q * g.nc[k];
        cell = cell * g.nc[k] + w;
        img |= ((q + 128) & 0xff) << (8 * k);
        sm.s_rel[lane][k] = (float)(pw[k] - (double)u[k] * g.h[k]);
      }
      sm.s_cell[lane] = cell;
      sm.s_img[lane] = img;
      end = (g.npbc == 3 || ((img >> 16) & 0xff) == 128) ? v.ccount[cell] : 0;
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int u = __shfl_up(end, o, kWave);
      if (lane >= o) end += u;
    }
    sm.s_end[lane] = end;
    const int total = __shfl(end, nloc - 1, kWave);
    wave_sync_lds();
    // candidates in rounds of kDescrLoads per lane: every load of a round is
    // issued before the first candidate of the round is used.  A lane's
    // candidates only grow, so its cell is found by walking on from the cell
    // of its previous candidate (a step or two: a cell holds tens of
    // candidates and the lane's next one is 64 nw further on) rather than by
    // a search over the chunk's prefix.
    int cur = 0;
    for (int r0 = 0; r0 < total; r0 += kDescrLoads * kWave * nw) {
      int tt[kDescrLoads], jj[kDescrLoads];
      float4 pp[kDescrLoads];
#pragma unroll
      for (int u = 0; u < kDescrLoads; u++) {
        const int it = r0 + (u * nw + wv) * kWave + lane;
        tt[u] = 0;
        jj[u] = -1;
        pp[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (it < total) {
          while (sm.s_end[cur] <= it) cur++;   // it < total = s_end[nloc - 1]: stops in the chunk
          const int slot = it - (cur ? sm.s_end[cur - 1] : 0);
          const size_t si = (size_t)sm.s_cell[cur] * g.cap + slot;
          tt[u] = cur;
          jj[u] = v.cbead[si];
          pp[u] = v.cpos32[si];
        }
      }
      // the round's candidates within r_d join the queue in (u, lane) order
      int qn = 0;
#pragma unroll
      for (int u = 0; u < kDescrLoads; u++) {
        const int j = jj[u], t = tt[u];
        const float4 pr = pp[u];
        const float fx = pr.x - sm.s_rel[t][0], fy = pr.y - sm.s_rel[t][1], fz = pr.z - sm.s_rel[t][2];
        const float r2 = fx * fx + fy * fy + fz * fz;
        const bool in = j >= 0 && !(contig ? (j >= mlo && j <= mhi) : v.moved[j]) && r2 < rd2;
        const unsigned long long mask = __ballot(in);
        if (in) {
          const int pos = qn + __popcll(mask & ((1ull << lane) - 1ull));
          const float r = sqrtf(r2);
          const float qq = qi * pr.w;
          sm.s_q[pos] = make_float2(r, 0.5f * (__cosf(pi_rd * r) + 1.0f));
          sm.s_qc[pos] = qq > 0.f ? 0 : (qq < 0.f ? 1 : 2);
        }
        qn += __popcll(mask);
      }
      wave_sync_lds();
#pragma unroll 4
      for (int e = 0; e < qn; e++) {
        const float2 q = sm.s_q[e];
        const float dr = q.x - my_mu;
        const float ph = __expf(-gam * dr * dr) * q.y;
        acc += sm.s_qc[e] == my_ch ? ph : 0.f;
      }
      wave_sync_lds();   // the next round rewrites the queue
    }
  }
  return acc;
}

// Feature f of the row in lane f from the summed basis values `sum`.
__device__ __forceinline__ float descr_finish(const Geo& g, const double4& p, float sum) {
  const int lane = threadIdx.x % kWave;
  if (lane < kRbf * kChan) return kFeatScale * sum;
  if (lane == kRbf * kChan) return (float)p.w;
  if (g.npbc == 2 && lane == kRbf * kChan + 1) return __expf(-0.5f * (float)p.z);
  if (g.npbc == 2 && lane == kRbf * kChan + 2) return __expf(-0.5f * (float)(g.Lsim[2] - p.z));
  return 0.f;
}

static_assert(kRbf * kChan + 3 <= kDescr && kDescr == kWave, "descriptor row = one value per lane");

// The moved beads' rows of the device loop.  System r's descriptor rows are
// 2 n_s per proposal s (old rows, then new rows) in proposal order; row j of
// them is bead m of proposal s at old (on = 0) or new position, physical row
// r sx_rows + (2 s + on) maxm + m.  A wave finds s from the inclusive prefix
// of 2 n_s over the proposals (one lane per proposal, the counts the proposal
// launch wrote: spec_rows).
__device__ __forceinline__ void spec_slot(int j, int pref, int np, int& s, int& on, int& m) {
  // s = number of proposals whose prefix ends at or before row j (every
  // lane runs the same six steps: the shuffles never diverge)
  s = 0;
#pragma unroll
  for (int step = kWave / 2; step > 0; step >>= 1) {
    const int cand = s + step;
    const int pv = __shfl(pref, min(cand, kWave) - 1, kWave);
    if (cand <= np && pv <= j) s = cand;
  }
  const int before = __shfl(pref, max(s - 1, 0), kWave);
  const int b0 = s ? before : 0;
  const int ns = (__shfl(pref, s, kWave) - b0) / 2;
  const int k = j - b0;
  on = k >= ns ? 1 : 0;
  m = k - on * ns;
}

__device__ __forceinline__ size_t spec_row(int j, int pref, int np, int r, int sx_rows, int maxm) {
  int s, on, m;
  spec_slot(j, pref, np, s, on, m);
  return (size_t)r * sx_rows + (size_t)(2 * s + on) * maxm + m;
}

// Rows of system d this batch (2 n_s summed over its proposals), the same in
// every lane; with `pref` each lane also gets the inclusive prefix up to
// proposal `lane` (the total in lanes past the last proposal).
__device__ __forceinline__ int spec_rows(const Dev& d, int lane, int* pref) {
  const int np = __builtin_amdgcn_readfirstlane(d.plan->n_prop);
  int p = lane < np ? 2 * d.recs[lane].n_moved : 0;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int u = __shfl_up(p, o, kWave);
    if (lane >= o) p += u;
  }
  if (pref) *pref = p;
  return __builtin_amdgcn_readfirstlane(__shfl(p, kWave - 1, kWave));
}

// Descriptors of every moved bead of every proposal, old and new: nbx blocks
// per system.  Each block takes kDescrWaves / wpr rows at a
// time, wpr waves per row splitting the row's candidates (a row holds
// hundreds to thousands of them: with few rows in a batch, one wave alone
// would wait on their loads in turn; with many, one wave per row saves the
// repeated cell setup); the partial sums of a row are added in wave order.
// Rows past the system's count are never visited, so a batch's grid does not
// grow with the padded row count.
__global__ __launch_bounds__(kBlock) void k_descr_spec(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                       const RepCfg* __restrict__ RC, int wpr, int R, int nbx) {
  __shared__ DescrWaveSmem sm[kDescrWaves];
  __shared__ float s_sum[kDescrWaves][kWave];
  // block (bx, r) of a grid of nbx blocks per system; when the systems split
  // evenly over the eight XCDs (workgroups are dispatched to XCD id % 8),
  // every block of system r runs on XCD r % 8, so each XCD's L2 holds the
  // cell lists of its own systems only
  int r, bx;
  if (R % kXcds == 0) {
    const int xcd = blockIdx.x % kXcds, k = blockIdx.x / kXcds, spx = R / kXcds;
    r = (k % spx) * kXcds + xcd;
    bx = k / spx;
  } else {
    r = blockIdx.x / nbx;
    bx = blockIdx.x - r * nbx;
  }
  const Geo& g = G[r];
  const Dev& d = D[r];
  const RepCfg& rc = RC[r];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int per = kDescrWaves / wpr, slot = w / wpr, part = w - slot * wpr;
  const int np = __builtin_amdgcn_readfirstlane(d.plan->n_prop);
  int pref;
  const int total = spec_rows(d, lane, &pref);
  for (int base = bx * per; base < total; base += nbx * per) {
    const int j = base + slot;
    float sum = 0.f;
    int s = 0, on = 0, m = 0;
    double4 p = make_double4(0, 0, 0, 0);
    if (j < total) {
      spec_slot(j, pref, np, s, on, m);
      const Dev v = prop_view(g, d, rc, s, 0);
      p = on ? v.mv_new[m] : v.mv_old[m];
      sum = descr_wave(g, v, rc.r_d, p, part, wpr, sm[w]);
    }
    s_sum[w][lane] = sum;
    __syncthreads();
    if (j < total && part == 0) {
#pragma unroll 1
      for (int k = 1; k < wpr; k++) sum += s_sum[w + k][lane];
      d.sur_x[((size_t)(2 * s + on) * rc.maxm + m) * kDescr + lane] = descr_finish(g, p, sum);
    }
    __syncthreads();   // the next rows rewrite s_sum
  }
}

// ------------------------------------------------------------ inference
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even.
__host__ __device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned int u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__host__ __device__ __forceinline__ float from_bf16(unsigned short h) {
  const unsigned int u = (unsigned int)h << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// Inference copy of the network: transposed BF16 weights (Wt[out][in]) and
// FP32 biases / output row.
struct NetBF16 {
  unsigned short w1t[kHid * kDescr];
  unsigned short w2t[kHid * kHid];
  unsigned short w3t[kHid * kHid];
  float b1[kHid], b2[kHid], b3[kHid], w4[kHid];
  float b4;
};

// LDS rows are padded by 4 BF16 (8 bytes): the 32 rows an MFMA operand load
// touches then start 2 banks apart instead of all in one bank.
constexpr int kPad = 4;
constexpr int kLd1 = kDescr + kPad, kLdH = kHid + kPad;
struct MlpSmem {
  __align__(16) unsigned short w1t[kHid * kLd1];
  __align__(16) unsigned short w2t[kHid * kLdH];
  __align__(16) unsigned short w3t[kHid * kLdH];
  float b1[kHid], b2[kHid], b3[kHid], w4[kHid];
  __align__(16) unsigned short act[kBlock / kWave][2][32 * kLdH];
};
static_assert(sizeof(MlpSmem) <= 160 * 1024, "fits the LDS of one CU");

// tanh of the inference layers: (1 - e) / (1 + e) with e = exp(-2 |x|), the
// sign of x restored.  Absolute error below 2e-7 (the rounding of e near
// x = 0), far under the BF16 rounding (2^-9 relative) applied to the result.
__device__ __forceinline__ float tanh_act(float x) {
  const float e = __expf(-2.0f * fabsf(x));
  return copysignf((1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e), x);
}

// One layer for the wave's 32 rows: out = tanh(in[32 x K] Wt^T + b) in BF16,
// column tiles nt0 .. nt0 + NT - 1 of 32 columns (LDS rows of in and Wt
// K + kPad apart, of out kHid + kPad).
// v_mfma_f32_32x32x8bf16_1k: lane l supplies A[l % 32][4 (l / 32) .. +3] and
// B[4 (l / 32) .. +3][l % 32]; it receives D[row][l % 32] for row =
// 8 (i / 4) + 4 (l / 32) + i % 4, i = 0..15.
template <int K, int NT>
__device__ __forceinline__ void mlp_layer(const unsigned short* in, const unsigned short* wt, const float* b,
                                          unsigned short* out, int nt0) {
  const int lane = threadIdx.x % kWave;
  const int r = lane % 32, hk = (lane / 32) * 4;
#pragma unroll
  for (int nt = nt0; nt < nt0 + NT; nt++) {
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.f;
    const int col = nt * 32 + r;
#pragma unroll
    for (int kk = 0; kk < K / 8; kk++) {
      bf16x4 a, bb;
      a = *reinterpret_cast<const bf16x4*>(in + r * (K + kPad) + kk * 8 + hk);
      bb = *reinterpret_cast<const bf16x4*>(wt + col * (K + kPad) + kk * 8 + hk);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int row = 8 * (i / 4) + hk + (i % 4);
      out[row * kLdH + col] = to_bf16(tanh_act(acc[i] + b[col]));
    }
  }
}

// Copy the BF16 network into LDS: every thread issues all its 16-byte loads
// (4 of W1, 8 each of W2 and W3) before its first LDS store, so the 80 KB
// arrive in one round trip rather than one per element.
__device__ __forceinline__ void stage_net(const NetBF16* __restrict__ net, MlpSmem& sm) {
  constexpr int n1 = kHid * kDescr / 8 / kBlock, n2 = kHid * kHid / 8 / kBlock;
  static_assert(n1 * 8 * kBlock == kHid * kDescr && n2 * 8 * kBlock == kHid * kHid, "whole 16-byte rows per thread");
  static_assert(offsetof(NetBF16, w2t) % 16 == 0 && offsetof(NetBF16, w3t) % 16 == 0, "16-byte weight rows");
  const int t = threadIdx.x;
  const uint4* g1 = reinterpret_cast<const uint4*>(net->w1t);
  const uint4* g2 = reinterpret_cast<const uint4*>(net->w2t);
  const uint4* g3 = reinterpret_cast<const uint4*>(net->w3t);
  uint4 r1[n1], r2[n2], r3[n2];
#pragma unroll
  for (int k = 0; k < n1; k++) r1[k] = g1[t + k * kBlock];
#pragma unroll
  for (int k = 0; k < n2; k++) {
    r2[k] = g2[t + k * kBlock];
    r3[k] = g3[t + k * kBlock];
  }
  float bb[4] = {0.f, 0.f, 0.f, 0.f};
  if (t < kHid) bb[0] = net->b1[t], bb[1] = net->b2[t], bb[2] = net->b3[t], bb[3] = net->w4[t];
  // 16 bytes of a global row land as two 8-byte halves of a padded LDS row
  auto put = [](unsigned short* dst, int ld, int per_row, int c, const uint4& v) {
    uint2* o = reinterpret_cast<uint2*>(dst + (c / per_row) * ld + (c % per_row) * 8);
    o[0] = make_uint2(v.x, v.y);
    o[1] = make_uint2(v.z, v.w);
  };
#pragma unroll
  for (int k = 0; k < n1; k++) put(sm.w1t, kLd1, kDescr / 8, t + k * kBlock, r1[k]);
#pragma unroll
  for (int k = 0; k < n2; k++) {
    put(sm.w2t, kLdH, kHid / 8, t + k * kBlock, r2[k]);
    put(sm.w3t, kLdH, kHid / 8, t + k * kBlock, r3[k]);
  }
  if (t < kHid) sm.b1[t] = bb[0], sm.b2[t] = bb[1], sm.b3[t] = bb[2], sm.w4[t] = bb[3];
  __syncthreads();
}

// f(row) for `rows` descriptor rows; each wave takes 32 rows at a time.
// Logical row i of `rows` is physical row (i / per) * stride + i % per of x
// and f (per = stride: contiguous rows).
__global__ __launch_bounds__(kBlock) void k_mlp_infer(const NetBF16* __restrict__ net, const float* __restrict__ x,
                                                      int rows, float* __restrict__ f, int per, int stride) {
  __shared__ MlpSmem sm;
  stage_net(net, sm);
  const float b4 = net->b4;
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  unsigned short* a0 = sm.act[w][0];
  unsigned short* a1 = sm.act[w][1];
  for (int base = (blockIdx.x * (kBlock / kWave) + w) * 32; base < rows; base += gridDim.x * (kBlock / kWave) * 32) {
    // input rows (zero past the end) in BF16, one row per pass; all 32 loads
    // are issued before the first store
    static_assert(kDescr == kWave, "one descriptor row per pass");
    float xv[32];
#pragma unroll
    for (int it = 0; it < 32; it++) {
      const int rr = base + it;
      const size_t pr = (size_t)(rr / per) * stride + rr % per;
      xv[it] = rr < rows ? x[pr * kDescr + lane] : 0.f;
    }
#pragma unroll
    for (int it = 0; it < 32; it++) a0[it * kLd1 + lane] = to_bf16(xv[it]);
    __builtin_amdgcn_wave_barrier();
    mlp_layer<kDescr, kHid / 32>(a0, sm.w1t, sm.b1, a1, 0);
    __builtin_amdgcn_wave_barrier();
    mlp_layer<kHid, kHid / 32>(a1, sm.w2t, sm.b2, a0, 0);
    __builtin_amdgcn_wave_barrier();
    mlp_layer<kHid, kHid / 32>(a0, sm.w3t, sm.b3, a1, 0);
    __builtin_amdgcn_wave_barrier();
    // output row: lanes r and r + 32 each take half of row r
    const int rrow = lane % 32, half = lane / 32;
    float s = 0;
    for (int j = half * (kHid / 2); j < (half + 1) * (kHid / 2); j++) s += from_bf16(a1[rrow * kLdH + j]) * sm.w4[j];
    s += __shfl_down(s, 32, kWave);
    const int ro = base + rrow;
    if (half == 0 && ro < rows) f[(size_t)(ro / per) * stride + ro % per] = s + b4;
    __builtin_amdgcn_wave_barrier();
  }
}

// Inference of the device loop over the moved beads' rows only: grid
// (blocks per system, systems); tile t of 32 rows (spec_row) of system r goes
// to block (t % gridDim.x, r), whose four waves share it (a quarter of the
// input rows, a quarter of each layer's columns and of the output rows each),
// so the few tiles of a batch run on different CUs with a short chain each;
// blocks past a system's tiles exit at once.
__global__ __launch_bounds__(kBlock) void k_mlp_spec(const NetBF16* __restrict__ net, const float* __restrict__ x,
                                                     float* __restrict__ f, const Dev* __restrict__ D,
                                                     const RepCfg* __restrict__ RC, int sx_rows) {
  static_assert(kBlock / kWave == kHid / 32, "one 32-column slice of each layer per wave");
  __shared__ MlpSmem sm;
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int r = blockIdx.y;
  const Dev& d = D[r];
  const int np = __builtin_amdgcn_readfirstlane(d.plan->n_prop);
  const int maxm = RC[r].maxm;
  int pref;
  const int total = spec_rows(d, lane, &pref);
  // a block with no tile of rows exits before staging the network
  if ((int)blockIdx.x * 32 >= total) return;
  stage_net(net, sm);
  const float b4 = net->b4;
  unsigned short* a0 = sm.act[0][0];
  unsigned short* a1 = sm.act[0][1];
  for (int lt = blockIdx.x; lt * 32 < total; lt += gridDim.x) {
    // lane q % 32 holds the physical row of tile row q (rows past the
    // system's end repeat its last row and are not written back)
    const int prow = (int)spec_row(min(lt * 32 + lane % 32, total - 1), pref, np, r, sx_rows, maxm);
    // wave w stages tile rows 8 w .. 8 w + 7 in BF16 (zero past the system's
    // rows), all eight loads issued before the first store
    float xv[8];
#pragma unroll
    for (int it = 0; it < 8; it++) {
      const int q = 8 * w + it;
      const int pr = __shfl(prow, q, kWave);
      xv[it] = lt * 32 + q < total ? x[(size_t)pr * kDescr + lane] : 0.f;
    }
#pragma unroll
    for (int it = 0; it < 8; it++) a0[(8 * w + it) * kLd1 + lane] = to_bf16(xv[it]);
    // each layer: wave w computes columns 32 w .. 32 w + 31 of all 32 rows
    __syncthreads();
    mlp_layer<kDescr, 1>(a0, sm.w1t, sm.b1, a1, w);
    __syncthreads();
    mlp_layer<kHid, 1>(a1, sm.w2t, sm.b2, a0, w);
    __syncthreads();
    mlp_layer<kHid, 1>(a0, sm.w3t, sm.b3, a1, w);
    __syncthreads();
    // output: wave w takes rows 8 w .. 8 w + 7, eight lanes of 16 columns
    // per row, summed in a fixed order
    const int q = 8 * w + lane / 8, part = lane % 8;
    float sum = 0;
#pragma unroll
    for (int j = part * 16; j < part * 16 + 16; j++) sum += from_bf16(a1[q * kLdH + j]) * sm.w4[j];
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) sum += __shfl_down(sum, o, 8);
    const int pq = __shfl(prow, q, kWave);
    if (part == 0 && lt * 32 + q < total) f[pq] = sum + b4;
    __syncthreads();   // the next tile rewrites both activation buffers
  }
}

// ------------------------------------------------------------ training (FP32)
// C[M x N] = act(A[M x K] B[K x N] + bias), act 0 = identity, 1 = tanh.
__global__ void k_gemm_ab(const float* __restrict__ A, const float* __restrict__ B, const float* __restrict__ bias,
                          float* __restrict__ Cm, int M, int N, int K, int act) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int i = idx / N, j = idx - (idx / N) * N;
  float s = bias ? bias[j] : 0.f;
  for (int k = 0; k < K; k++) s += A[(size_t)i * K + k] * B[(size_t)k * N + j];
  Cm[idx] = act ? tanhf(s) : s;
}

// C[M x N] = (A[M x K] B^T) * (1 - H^2), B given as [N x K]... here B is the
// weight W[N x K] (rows = layer input), so C = A W^T elementwise times the
// tanh derivative of H[M x N].
__global__ void k_gemm_abt_dtanh(const float* __restrict__ A, const float* __restrict__ W,
                                 const float* __restrict__ H, float* __restrict__ Cm, int M, int N, int K) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int i = idx / N, j = idx - (idx / N) * N;
  float s = 0.f;
  for (int k = 0; k < K; k++) s += A[(size_t)i * K + k] * W[(size_t)j * K + k];
  const float h = H[idx];
  Cm[idx] = s * (1.f - h * h);
}

// G[K1 x K2] = A^T B for A[M x K1], B[M x K2] (weight gradients) in two
// passes: k_gemm_atb_part sums rows [c R, (c + 1) R) of chunk c into
// part[c][K1 x K2 (+ K2 column sums of B)], k_gemm_atb_sum adds the chunks in
// order.
constexpr int kAtbRows = 128;

__global__ void k_gemm_atb_part(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ part,
                                int M, int K1, int K2) {
  const int c = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nout = K1 * K2 + K2;
  if (idx >= nout) return;
  const int m0 = c * kAtbRows, m1 = min(M, m0 + kAtbRows);
  float s = 0.f;
  if (idx < K1 * K2) {
    const int i = idx / K2, j = idx - (idx / K2) * K2;
    for (int m = m0; m < m1; m++) s += A[(size_t)m * K1 + i] * B[(size_t)m * K2 + j];
  } else {
    const int j = idx - K1 * K2;
    for (int m = m0; m < m1; m++) s += B[(size_t)m * K2 + j];
  }
  part[(size_t)c * nout + idx] = s;
}

__global__ void k_gemm_atb_sum(const float* __restrict__ part, int chunks, int K1, int K2, float* __restrict__ Gm,
                               float* __restrict__ bias_g) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nout = K1 * K2 + K2;
  if (idx >= nout) return;
  float s = 0.f;
  for (int c = 0; c < chunks; c++) s += part[(size_t)c * nout + idx];
  if (idx < K1 * K2) Gm[idx] = s;
  else if (bias_g) bias_g[idx - K1 * K2] = s;
}

// Minibatch loss: sample i owns rows [i R, (i + 1) R) of f (R = 2 maxm: old
// beads then new beads, n_i of each in use).  pred_i = sum(new) - sum(old);
// loss = mean (pred_i - clip(target_i))^2; g[row] = dloss/df(row).
__global__ void k_train_loss(const float* __restrict__ f, const double* __restrict__ target,
                             const int* __restrict__ n, int samples, int maxm, double clip, float* __restrict__ g,
                             double* __restrict__ loss_out) {
  __shared__ double sh[kBlock / kWave];
  double l = 0;
  for (int i = threadIdx.x; i < samples; i += kBlock) {
    const float* fo = f + (size_t)i * 2 * maxm;
    const float* fn = fo + maxm;
    double pred = 0;
    for (int m = 0; m < n[i]; m++) pred += (double)fn[m] - (double)fo[m];
    const double tgt = fmin(fmax(target[i], -clip), clip);
    const double res = pred - tgt;
    l += res * res;
    const float gr = (float)(2.0 * res / samples);
    float* go = g + (size_t)i * 2 * maxm;
    for (int m = 0; m < maxm; m++) {
      go[m] = m < n[i] ? -gr : 0.f;
      go[maxm + m] = m < n[i] ? gr : 0.f;
    }
  }
  const double t = block_sum<kBlock>(l, sh);
  if (threadIdx.x == 0) *loss_out = t / samples;
}

// delta3[M x 128] = g[m] w4 (1 - H3^2).
__global__ void k_out_delta(const float* __restrict__ g, const float* __restrict__ w4, const float* __restrict__ H3,
                            float* __restrict__ d3, int M) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * kHid) return;
  const int m = idx / kHid, j = idx - (idx / kHid) * kHid;
  const float h = H3[idx];
  d3[idx] = g[m] * w4[j] * (1.f - h * h);
}

// Adam (Kingma & Ba 2015) on the FP32 master copy, then the BF16 inference
// copy (transposed).  step = 1-based step count.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = grad[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float ah = a / (1.f - powf(b1, (float)step));
  const float bh = b / (1.f - powf(b2, (float)step));
  w[i] -= lr * ah / (sqrtf(bh) + eps);
}

__global__ void k_pack_bf16(const float* __restrict__ w, NetBF16* __restrict__ net) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kHid * kDescr) {
    const int o = i / kDescr, in = i - (i / kDescr) * kDescr;
    net->w1t[i] = to_bf16(w[kOffW1 + in * kHid + o]);
  }
  if (i < kHid * kHid) {
    const int o = i / kHid, in = i - (i / kHid) * kHid;
    net->w2t[i] = to_bf16(w[kOffW2 + in * kHid + o]);
    net->w3t[i] = to_bf16(w[kOffW3 + in * kHid + o]);
  }
  if (i < kHid) {
    net->b1[i] = w[kOffB1 + i];
    net->b2[i] = w[kOffB2 + i];
    net->b3[i] = w[kOffB3 + i];
    net->w4[i] = w[kOffW4 + i];
  }
  if (i == 0) net->b4 = w[kOffB4];
}

// Minibatch gather: the `samples` most recent ring slots (newest last).
__global__ void k_gather_samples(const float* __restrict__ tr_x, const double* __restrict__ tr_target,
                                 const int* __restrict__ tr_n, const int* __restrict__ tr_ctl, int samples,
                                 int maxm, float* __restrict__ X, double* __restrict__ target, int* __restrict__ n) {
  const int i = blockIdx.y;
  if (i >= samples) return;
  const int head = tr_ctl[0], ring = tr_ctl[1];
  const int slot = ((head - samples + i) % ring + ring) % ring;
  const size_t rowsz = (size_t)2 * maxm * kDescr;
  for (size_t k = blockIdx.x * blockDim.x + threadIdx.x; k < rowsz; k += (size_t)gridDim.x * blockDim.x)
    X[(size_t)i * rowsz + k] = tr_x[(size_t)slot * rowsz + k];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    target[i] = tr_target[slot];
    n[i] = tr_n[slot];
  }
}

}  // namespace plum
