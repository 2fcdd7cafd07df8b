// Batched replicas: R independent systems advance together, every launch
// carrying one trial move of each.  A k_compute_batch launch spreads the
// virtual blocks of all replicas (k space, real-space queries, moved-moved
// pairs; the same block functions as the single-system kernel) over a fixed
// grid; k_serial_batch runs one block per replica (epilogue, commit, next
// proposal).  Kernel boundaries are the only synchronisation.
#pragma once
#include "kernels_move.h"

namespace plum {

constexpr int kMaxReplicas = 1024;
constexpr int kMaxSplit = 8;           // most blocks per real-space query
constexpr int kDefaultSplitDiv = 384;  // candidate slots per query block

// Per-replica launch shape.
struct RepCfg {
  int nkb;      // k-space blocks (0 without Ewald)
  int nsplit;   // blocks per real-space query
  int ncell;    // cells
  int pad;
};

// Moved-moved pair blocks: 256 pairs each, independent of the launch size, so
// the partial sums come out the same in every launch shape.
__host__ __device__ __forceinline__ int nmm_blocks(int nm) {
  const int npair = nm * (nm - 1) / 2;
  return npair > 0 ? (npair + kBlock - 1) / kBlock : 1;
}

__device__ __forceinline__ int virtual_blocks(const RepCfg& rc, int nm) {
  return rc.nkb + 2 * nm * rc.nsplit + nmm_blocks(nm);
}

// Replica r owning virtual block v: first r with pre[r] > v (inclusive prefix).
__device__ __forceinline__ int find_replica(const int* pre, int R, int v) {
  int lo = 0, hi = R - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] > v) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// The per-replica records are read through the constant address space and
// copied by value for each virtual block: the loads then stay scalar and
// rematerialisable across the block functions' barriers (through a generic
// pointer they turn into per-lane loads and double the register count).
template <typename T>
__device__ __forceinline__ T load_const(const T* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(4))) const T CT;
  return ((CT*)p)[i];
#else
  return p[i];
#endif
}

template <int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_compute_batch(
    const Geo* __restrict__ G, const Dev* __restrict__ D, const RepCfg* __restrict__ RC,
    const int* __restrict__ vb_count, int R) {
  __shared__ int s_pre[kMaxReplicas];
  __shared__ union U {
    KSmem k;
    RSmem<P> r;
  } sm;
  for (int r = threadIdx.x; r < R; r += kBlock) s_pre[r] = vb_count[r];
  __syncthreads();
  lds_scan(s_pre, R);
  const int total = s_pre[R - 1];
  for (int v = blockIdx.x; v < total; v += gridDim.x) {
    const int r = __builtin_amdgcn_readfirstlane(find_replica(s_pre, R, v));
    const int b = __builtin_amdgcn_readfirstlane(v - (r ? s_pre[r - 1] : 0));
    const Geo g = load_const(G, r);
    const Dev d = load_const(D, r);
    const RepCfg rc = load_const(RC, r);
    if (b < rc.nkb) {
      kspace_block(g, d, b, sm.k);
    } else {
      const int nm = d.rec->n_moved;
      real_block<P>(g, d, b - rc.nkb, 2 * nm * rc.nsplit, nmm_blocks(nm), rc.nsplit, sm.r);
    }
    __syncthreads();  // the LDS union is reused by the next virtual block
  }
}

// One block per replica: epilogue + Metropolis + commit of the pending move
// (do_epi), then the next proposal, its phase tables and its virtual-block
// count (do_prop).
__global__ __launch_bounds__(kBlock) void k_serial_batch(const Geo* __restrict__ G, const Dev* __restrict__ D,
                                                         const McDev* __restrict__ MC,
                                                         const RepCfg* __restrict__ RC, int* vb_count, int R,
                                                         int do_epi, int do_prop) {
  __shared__ union U {
    EpSmem e;
    PrSmem p;
  } sm;
  const int r = blockIdx.x;
  if (r >= R) return;
  const Geo& g = G[r];
  const Dev& d = D[r];
  const RepCfg& rc = RC[r];
  if (do_epi) {
    const int nm = d.rec->n_moved;
    epilogue_block(g, d, rc.nkb, 2 * nm * rc.nsplit, nmm_blocks(nm), 1, MC[r], sm.e);
  }
  __syncthreads();
  if (do_prop) {
    propose_block(g, d, MC[r], sm.p);
    __syncthreads();
    if (g.use_ewald && g.n_kh > 0) tables_block(g, d);
    if (threadIdx.x == 0) vb_count[r] = virtual_blocks(rc, d.rec->n_moved);
  }
}

// Tombstone compaction of every replica's cells; cell_off is the exclusive
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
  const int r = lo;
  const Geo& g = G[r];
  const Dev& d = D[r];
  compact_cell(g, d, gid - cell_off[r]);
}

// Overflow flags of all replicas into one array (read by the host between
// graph replays).
__global__ void k_batch_flags(const Dev* __restrict__ D, int R, int* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) out[r] = D[r].ctl->overflow;
}

// Pending S(k) += dS(k) of every replica (end of a run); one block row per
// replica, grid-stride over its k vectors.
__global__ void k_flush_sk_batch(const Geo* __restrict__ G, const Dev* __restrict__ D, int R) {
  const int r = blockIdx.y;
  if (r >= R) return;
  const Geo& g = G[r];
  const Dev& d = D[r];
  if (!d.ctl->sk_pending) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < g.n_kh; k += gridDim.x * blockDim.x) {
    double2 s = d.Sk[k];
    const double2 ds = d.dS[k];
    s.x += ds.x;
    s.y += ds.y;
    d.Sk[k] = s;
  }
}

__global__ void k_clear_sk_pending_batch(const Dev* __restrict__ D, int R) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) D[r].ctl->sk_pending = 0;
}

}  // namespace plum
