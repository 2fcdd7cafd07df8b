// plum_cbmc_trials_noalloc (include/plum_gpu.h): the configurational-bias
// trial energies of plum_cbmc_trials with no device allocation.  Included
// once, at the end of engine.hip, after cbmc_api.h.
// ---- plum_cbmc_trials without allocation ---------------------------------
// The variant for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops): the request, the trial and grown-bead
// coordinates, the partial sums and the energies are staged in one
// file-scope device array instead of context buffers allocated on first use.
// Every call on a device shares that device's array, so calls that overlap on
// one device race; nothing in it has to start at zero (every partial slot a
// call reads is written by that call first).
namespace {

constexpr int kNaMaxTrials = 64;      // K
constexpr int kNaMaxGrown = 512;      // chain beads grown so far
constexpr int kNaMaxKBlocks = 128;    // k-space blocks of 256 vectors (n_kh <= 32768)
constexpr size_t kNaOffIn = 0;                                          // CbmcIn (64 B)
constexpr size_t kNaOffXyz = 8;                                         // 6 K doubles
constexpr size_t kNaOffGrown = kNaOffXyz + 6 * kNaMaxTrials;            // 6 cur doubles
constexpr size_t kNaOffCells = kNaOffGrown + 6 * kNaMaxGrown;           // 2 K kMaxSplit double2
constexpr size_t kNaOffKpart = kNaOffCells + 2 * 2 * kNaMaxTrials * kMaxSplit;  // K nkb doubles
constexpr size_t kNaOffOut = kNaOffKpart + (size_t)kNaMaxTrials * kNaMaxKBlocks;  // K doubles
constexpr size_t kNaOffMz = kNaOffOut + kNaMaxTrials;                   // slab dipole M_z, 1 double
constexpr size_t kNaDoubles = kNaOffMz + 1;
static_assert(sizeof(CbmcIn) <= 8 * sizeof(double), "request slot");
static_assert(kNaDoubles <= 16384, "scratch stays at 128 KiB");

}  // namespace

static __device__ __align__(256) double scratch[16384];

namespace {

std::mutex g_na_mu;
double* g_na_ptr[64] = {nullptr};

// The device address of `scratch` on the context's device, looked up once
// per device with that device current.
int na_scratch(plum_ctx* c, double** out) {
  if (c->device < 0 || c->device >= 64) return fail(PLUM_E_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[c->device]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != c->device) HIPCHK(hipSetDevice(c->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != c->device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(PLUM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[c->device] = static_cast<double*>(p);
  }
  *out = g_na_ptr[c->device];
  return PLUM_OK;
}

}  // namespace

extern "C" int plum_cbmc_trials_noalloc(plum_ctx* c, const plum_cbmc_request* rq, const double* trials,
                                        const double* grown, double* energies) {
  if (!c || !rq || !trials || !energies) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  const int K = rq->n_trials, cur = rq->grown, ch = rq->charged ? 1 : 0;
  if (K < 1 || K > kNaMaxTrials) return fail(PLUM_E_CAPACITY, "n_trials must be 1..64 without allocation");
  if (cur < 0 || cur > kNaMaxGrown) return fail(PLUM_E_CAPACITY, "grown must be 0..512 without allocation");
  if (cur > 0 && !grown) return fail(PLUM_E_ARG, "null grown beads");
  const Geo& g = c->g;
  const int nkb = (g.use_ewald && g.n_kh > 0) ? (g.n_kh + kBlock - 1) / kBlock : 0;
  if (nkb > kNaMaxKBlocks) return fail(PLUM_E_CAPACITY, "more than 32768 k vectors without allocation");
  if (c->nsplit > kMaxSplit) return fail(PLUM_E_CAPACITY, "real-space split beyond the scratch layout");
  if (rq->type_chain < 0 || rq->type_chain >= g.n_types || (ch && (rq->type_ion < 0 || rq->type_ion >= g.n_types)))
    return fail(PLUM_E_ARG, "bead type out of range");
  const bool ex = rq->exclude_lo <= rq->exclude_hi;
  if (ex && (rq->exclude_lo < 0 || rq->exclude_hi >= g.n)) return fail(PLUM_E_ARG, "exclusion range out of range");
  for (int i = 0; i < 3 * K * (1 + ch); i++)
    if (!std::isfinite(trials[i])) return fail(PLUM_E_ARG, "non-finite trial coordinate");
  double* sp = nullptr;
  CHK(na_scratch(c, &sp));
  CbmcIn in{};
  in.K = K;
  in.cur = cur;
  in.ex_lo = ex ? rq->exclude_lo : 0;
  in.ex_hi = ex ? rq->exclude_hi : -1;
  in.type_c = rq->type_chain;
  in.type_i = ch ? rq->type_ion : rq->type_chain;
  in.charged = ch;
  in.q_c = rq->q_chain;
  in.q_i = ch ? rq->q_ion : 0.0;
  CbmcIn* d_in = reinterpret_cast<CbmcIn*>(sp + kNaOffIn);
  double* d_xyz = sp + kNaOffXyz;
  double* d_grown = sp + kNaOffGrown;
  double2* d_cells = reinterpret_cast<double2*>(sp + kNaOffCells);
  double* d_kpart = sp + kNaOffKpart;
  double* d_out = sp + kNaOffOut;
  double* d_mz = sp + kNaOffMz;
  // work still queued by the device loop comes first
  if (c->self) {
    HIPCHK(hipEventRecord(c->ev_idle, c->self->st));
    HIPCHK(hipStreamWaitEvent(c->st, c->ev_idle, 0));
  }
  HIPCHK(hipMemcpyAsync(d_in, &in, sizeof(CbmcIn), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(d_xyz, trials, sizeof(double) * 3 * K * (1 + ch), hipMemcpyHostToDevice, c->st));
  if (cur > 0)
    HIPCHK(hipMemcpyAsync(d_grown, grown, sizeof(double) * 3 * cur * (1 + ch), hipMemcpyHostToDevice, c->st));
  const int ns = c->nsplit;
  if (c->p.precision == PLUM_FP32_PAIR)
    hipLaunchKernelGGL(k_cbmc_cells<1>, dim3(2 * K * ns), dim3(kBlock), 0, c->st, g, c->d, d_in, d_xyz, ns, d_cells);
  else
    hipLaunchKernelGGL(k_cbmc_cells<0>, dim3(2 * K * ns), dim3(kBlock), 0, c->st, g, c->d, d_in, d_xyz, ns, d_cells);
  if (nkb > 0)
    hipLaunchKernelGGL(k_cbmc_kspace, dim3(nkb), dim3(kBlock), 0, c->st, g, c->d, d_in, d_xyz, d_grown, d_kpart);
  if (g.dipole && g.use_ewald)
    hipLaunchKernelGGL(k_cbmc_mz, dim3(1), dim3(kBlock), 0, c->st, g, c->d, d_in, d_grown, d_mz);
  hipLaunchKernelGGL(k_cbmc_final, dim3(K), dim3(kBlock), 0, c->st, g, c->d, d_in, d_xyz, d_grown, d_cells, ns,
                     d_kpart, nkb, d_mz, d_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(energies, d_out, sizeof(double) * K, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return PLUM_OK;
}
