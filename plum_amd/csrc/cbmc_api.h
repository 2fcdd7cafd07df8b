// Host side of plum_cbmc_trials (include/plum_gpu.h).  Included once, at the
// end of engine.hip.

struct CbmcHost {
  int cap_k = 0, cap_g = 0, nkb = 0;
  CbmcIn* in = nullptr;
  double *xyz = nullptr, *grown = nullptr, *kpart = nullptr, *out = nullptr;
  double2* cells = nullptr;
  char* h_stage = nullptr;   // pinned: CbmcIn + trials + grown
  double* h_out = nullptr;   // pinned
};

void cbmc_free(CbmcHost* h) {
  if (!h) return;
  if (h->h_stage) (void)hipHostFree(h->h_stage);
  if (h->h_out) (void)hipHostFree(h->h_out);
  delete h;
}

namespace {

constexpr int kCbmcMaxTrials = 1024;

int cbmc_reserve(plum_ctx* c, int K, int cur) {
  CbmcHost*& h = c->cbmc;
  if (!h) h = new CbmcHost();
  if (K <= h->cap_k && cur <= h->cap_g) return PLUM_OK;
  const int ck = std::max(K, h->cap_k), cg = std::max(cur, h->cap_g);
  h->nkb = std::max(1, (c->g.n_kh + kBlock - 1) / kBlock);
  CHK(dalloc(c, &h->in, 1));
  CHK(dalloc(c, &h->xyz, (size_t)6 * ck));
  CHK(dalloc(c, &h->grown, (size_t)6 * std::max(1, cg)));
  CHK(dalloc(c, &h->cells, (size_t)2 * ck * kMaxSplit));
  CHK(dalloc(c, &h->kpart, (size_t)ck * h->nkb));
  CHK(dalloc(c, &h->out, ck));
  if (h->h_stage) (void)hipHostFree(h->h_stage);
  if (h->h_out) (void)hipHostFree(h->h_out);
  h->h_stage = nullptr;
  h->h_out = nullptr;
  const size_t stage = sizeof(CbmcIn) + sizeof(double) * (6 * (size_t)ck + 6 * (size_t)std::max(1, cg));
  if (hipHostMalloc((void**)&h->h_stage, stage, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&h->h_out, sizeof(double) * ck, hipHostMallocDefault) != hipSuccess)
    return fail(PLUM_E_NOMEM, "hipHostMalloc failed");
  h->cap_k = ck;
  h->cap_g = cg;
  return PLUM_OK;
}

}  // namespace

extern "C" int plum_cbmc_trials(plum_ctx* c, const plum_cbmc_request* rq, const double* trials,
                                const double* grown, double* energies) {
  if (!c || !rq || !trials || !energies) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  const int K = rq->n_trials, cur = rq->grown, ch = rq->charged ? 1 : 0;
  if (K < 1 || K > kCbmcMaxTrials) return fail(PLUM_E_ARG, "n_trials must be 1..1024");
  if (cur < 0 || cur > 2 * kMaxChainDev) return fail(PLUM_E_ARG, "grown count out of range");
  if (cur > 0 && !grown) return fail(PLUM_E_ARG, "null grown beads");
  if (rq->type_chain < 0 || rq->type_chain >= c->g.n_types || (ch && (rq->type_ion < 0 || rq->type_ion >= c->g.n_types)))
    return fail(PLUM_E_ARG, "bead type out of range");
  const bool ex = rq->exclude_lo <= rq->exclude_hi;
  if (ex && (rq->exclude_lo < 0 || rq->exclude_hi >= c->g.n)) return fail(PLUM_E_ARG, "exclusion range out of range");
  if (c->g.dipole) return fail(PLUM_E_UNSUPPORTED, "configurational-bias trials with the slab dipole term");
  for (int i = 0; i < 3 * K * (1 + ch); i++)
    if (!std::isfinite(trials[i])) return fail(PLUM_E_ARG, "non-finite trial coordinate");
  CHK(cbmc_reserve(c, K, cur));
  CbmcHost* h = c->cbmc;
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
  std::memcpy(h->h_stage, &in, sizeof(in));
  double* st_xyz = reinterpret_cast<double*>(h->h_stage + sizeof(CbmcIn));
  std::memcpy(st_xyz, trials, sizeof(double) * 3 * K * (1 + ch));
  if (!ch) std::memset(st_xyz + 3 * K, 0, sizeof(double) * 3 * K);
  double* st_g = st_xyz + 6 * (size_t)K;
  if (cur > 0) std::memcpy(st_g, grown, sizeof(double) * 3 * cur * (1 + ch));
  HIPCHK(hipMemcpyAsync(h->in, h->h_stage, sizeof(CbmcIn), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(h->xyz, st_xyz, sizeof(double) * 6 * K, hipMemcpyHostToDevice, c->st));
  if (cur > 0)
    HIPCHK(hipMemcpyAsync(h->grown, st_g, sizeof(double) * 3 * cur * (1 + ch), hipMemcpyHostToDevice, c->st));
  const Geo& g = c->g;
  const int ns = c->nsplit;
  if (c->p.precision == PLUM_FP32_PAIR)
    hipLaunchKernelGGL(k_cbmc_cells<1>, dim3(2 * K * ns), dim3(kBlock), 0, c->st, g, c->d, h->in, h->xyz, ns,
                       h->cells);
  else
    hipLaunchKernelGGL(k_cbmc_cells<0>, dim3(2 * K * ns), dim3(kBlock), 0, c->st, g, c->d, h->in, h->xyz, ns,
                       h->cells);
  const int nkb = (g.use_ewald && g.n_kh > 0) ? h->nkb : 0;
  if (nkb > 0)
    hipLaunchKernelGGL(k_cbmc_kspace, dim3(nkb), dim3(kBlock), 0, c->st, g, c->d, h->in, h->xyz, h->grown, h->kpart);
  hipLaunchKernelGGL(k_cbmc_final, dim3(K), dim3(kBlock), 0, c->st, g, c->d, h->in, h->xyz, h->grown, h->cells, ns,
                     h->kpart, nkb, h->out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->h_out, h->out, sizeof(double) * K, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  std::memcpy(energies, h->h_out, sizeof(double) * K);
  return PLUM_OK;
}
