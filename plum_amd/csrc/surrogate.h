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
// Per-wave scratch of descr_wave: one chunk of up to 64 query cells.
struct DescrWaveSmem {
  int s_cell[kWave];
  int s_img[kWave];
  int s_end[kWave];
  float s_rel[kWave][3];
};
constexpr int kDescrWaves = kBlock / kWave;

// Descriptor of moved bead m of proposal view v at its old (on = 0) or new
// (on = 1) position, against every unmoved bead image within r_d (images as
// the LJ term takes them: z unshifted with two periodic dimensions).  One
// wave per row, no block barriers: the wave scans the query cells 64 at a
// time, each lane sums its candidates, and a butterfly over the 64 lanes
// (32 + 16 + ... + 1 exchanges, fixed order) leaves feature f in lane f.
__device__ void descr_wave(const Geo& g, const Dev& v, float r_d, int m, int on, float* out, DescrWaveSmem& sm) {
  const int lane = threadIdx.x % kWave;
  const double4 p = on ? v.mv_new[m] : v.mv_old[m];
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
  float acc[kWave];
#pragma unroll
  for (int f = 0; f < kWave; f++) acc[f] = 0.f;
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
    for (int it = lane; it < total; it += kWave) {
      int lo_i = 0, hi_i = nloc - 1;
      while (lo_i < hi_i) {
        const int mid = (lo_i + hi_i) >> 1;
        if (sm.s_end[mid] > it) hi_i = mid;
        else lo_i = mid + 1;
      }
      const int t = lo_i;
      const int slot = it - (t ? sm.s_end[t - 1] : 0);
      const size_t si = (size_t)sm.s_cell[t] * g.cap + slot;
      const int j = v.cbead[si];
      if (j < 0 || (contig ? (j >= mlo && j <= mhi) : v.moved[j])) continue;
      const float4 pr = v.cpos32[si];
      const float fx = pr.x - sm.s_rel[t][0], fy = pr.y - sm.s_rel[t][1], fz = pr.z - sm.s_rel[t][2];
      const float r2 = fx * fx + fy * fy + fz * fz;
      if (r2 >= rd2) continue;
      const float r = sqrtf(r2);
      const float fc = 0.5f * (__cosf(pi_rd * r) + 1.0f);
      const float qq = qi * pr.w;
      const int ch = qq > 0.f ? 0 : (qq < 0.f ? 1 : 2);
#pragma unroll
      for (int k = 0; k < kRbf; k++) {
        const float dr = r - (kRbfMin + k * dmu);
        const float ph = __expf(-gam * dr * dr) * fc;
        acc[k] += ch == 0 ? ph : 0.f;
        acc[kRbf + k] += ch == 1 ? ph : 0.f;
        acc[2 * kRbf + k] += ch == 2 ? ph : 0.f;
      }
    }
    wave_sync_lds();   // the next chunk rewrites the cell arrays
  }
  // butterfly: at offset o each lane keeps the half of its 2o values that its
  // bit o selects and adds its partner's copy of that half
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < o; i++) {
      const float keep = up ? acc[i + o] : acc[i];
      const float give = up ? acc[i] : acc[i + o];
      acc[i] = keep + __shfl_xor(give, o, kWave);
    }
  }
  float val;
  if (lane < kRbf * kChan) val = kFeatScale * acc[0];
  else if (lane == kRbf * kChan) val = qi;
  else if (g.npbc == 2 && lane == kRbf * kChan + 1) val = __expf(-0.5f * (float)p.z);
  else if (g.npbc == 2 && lane == kRbf * kChan + 2) val = __expf(-0.5f * (float)(g.Lsim[2] - p.z));
  else val = 0.f;
  if (lane < kDescr) out[lane] = val;
}

static_assert(kRbf * kChan + 3 <= kDescr && kDescr == kWave, "descriptor row = one value per lane");

// Descriptors of every moved bead of every proposal, old and new: grid
// (2 maxm / 4, pmax, systems), one wave per row; row m2 of system r is bead
// m2 % maxm at old (m2 < maxm) or new.
__global__ __launch_bounds__(kBlock) void k_descr_spec(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                       const RepCfg* __restrict__ RC) {
  __shared__ DescrWaveSmem sm[kDescrWaves];
  const Geo& g = G[blockIdx.z];
  const Dev& d = D[blockIdx.z];
  const RepCfg& rc = RC[blockIdx.z];
  const int w = threadIdx.x / kWave;
  const int m2 = blockIdx.x * kDescrWaves + w;
  if (m2 >= 2 * rc.maxm) return;
  const int s = blockIdx.y;
  if (s >= d.plan->n_prop) return;
  const int on = m2 >= rc.maxm ? 1 : 0, m = m2 - on * rc.maxm;
  const Dev v = prop_view(g, d, rc, s, 0);
  if (m >= v.rec->n_moved) return;
  descr_wave(g, v, rc.r_d, m, on, d.sur_x + ((size_t)(2 * s + on) * rc.maxm + m) * kDescr, sm[w]);
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

struct MlpSmem {
  __align__(16) unsigned short w1t[kHid * kDescr];
  __align__(16) unsigned short w2t[kHid * kHid];
  __align__(16) unsigned short w3t[kHid * kHid];
  float b1[kHid], b2[kHid], b3[kHid], w4[kHid];
  __align__(16) unsigned short act[kBlock / kWave][2][32 * kHid];
};

// One layer for the wave's 32 rows: out = tanh(in[32 x K] Wt^T + b) in BF16.
// v_mfma_f32_32x32x8bf16_1k: lane l supplies A[l % 32][4 (l / 32) .. +3] and
// B[4 (l / 32) .. +3][l % 32]; it receives D[row][l % 32] for row =
// 8 (i / 4) + 4 (l / 32) + i % 4, i = 0..15.
template <int K>
__device__ __forceinline__ void mlp_layer(const unsigned short* in, const unsigned short* wt, const float* b,
                                          unsigned short* out) {
  const int lane = threadIdx.x % kWave;
  const int r = lane % 32, hk = (lane / 32) * 4;
#pragma unroll
  for (int nt = 0; nt < kHid / 32; nt++) {
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.f;
    const int col = nt * 32 + r;
#pragma unroll
    for (int kk = 0; kk < K / 8; kk++) {
      bf16x4 a, bb;
      a = *reinterpret_cast<const bf16x4*>(in + r * K + kk * 8 + hk);
      bb = *reinterpret_cast<const bf16x4*>(wt + col * K + kk * 8 + hk);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int row = 8 * (i / 4) + hk + (i % 4);
      out[row * kHid + col] = to_bf16(tanhf(acc[i] + b[col]));
    }
  }
}

// f(row) for `rows` descriptor rows; each wave takes 32 rows at a time.
// Logical row i of `rows` is physical row (i / per) * stride + i % per of x
// and f (per = stride: contiguous rows).
__global__ __launch_bounds__(kBlock) void k_mlp_infer(const NetBF16* __restrict__ net, const float* __restrict__ x,
                                                      int rows, float* __restrict__ f, int per, int stride) {
  __shared__ MlpSmem sm;
  for (int i = threadIdx.x; i < kHid * kDescr; i += kBlock) sm.w1t[i] = net->w1t[i];
  for (int i = threadIdx.x; i < kHid * kHid; i += kBlock) {
    sm.w2t[i] = net->w2t[i];
    sm.w3t[i] = net->w3t[i];
  }
  for (int i = threadIdx.x; i < kHid; i += kBlock) {
    sm.b1[i] = net->b1[i];
    sm.b2[i] = net->b2[i];
    sm.b3[i] = net->b3[i];
    sm.w4[i] = net->w4[i];
  }
  const float b4 = net->b4;
  __syncthreads();
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  unsigned short* a0 = sm.act[w][0];
  unsigned short* a1 = sm.act[w][1];
  for (int base = (blockIdx.x * (kBlock / kWave) + w) * 32; base < rows; base += gridDim.x * (kBlock / kWave) * 32) {
    // input rows (zero past the end) in BF16
    for (int i = lane; i < 32 * kDescr; i += kWave) {
      const int rr = base + i / kDescr;
      const size_t pr = (size_t)(rr / per) * stride + rr % per;
      a0[i] = to_bf16(rr < rows ? x[pr * kDescr + (i % kDescr)] : 0.f);
    }
    __builtin_amdgcn_wave_barrier();
    mlp_layer<kDescr>(a0, sm.w1t, sm.b1, a1);
    __builtin_amdgcn_wave_barrier();
    mlp_layer<kHid>(a1, sm.w2t, sm.b2, a0);
    __builtin_amdgcn_wave_barrier();
    mlp_layer<kHid>(a0, sm.w3t, sm.b3, a1);
    __builtin_amdgcn_wave_barrier();
    // output row: lanes r and r + 32 each take half of row r
    const int rrow = lane % 32, half = lane / 32;
    float s = 0;
    for (int j = half * (kHid / 2); j < (half + 1) * (kHid / 2); j++) s += from_bf16(a1[rrow * kHid + j]) * sm.w4[j];
    s += __shfl_down(s, 32, kWave);
    const int ro = base + rrow;
    if (half == 0 && ro < rows) f[(size_t)(ro / per) * stride + ro % per] = s + b4;
    __builtin_amdgcn_wave_barrier();
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
