// Host side of plum_cbmc_trials (include/plum_gpu.h).  Included once, at the
// end of engine.hip.

struct CbmcHost {
  // whole-chain growth: buffers for K trials x L steps, and the captured
  // launch sequence of one growth
  int gK = 0, gL = 0, g_nkb = 0, g_prec = -1;
  GrowIn* g_in = nullptr;
  GrowOut* g_out = nullptr;
  double *g_rnd = nullptr, *g_grown = nullptr, *g_kpart = nullptr;
  double2 *g_Sx = nullptr, *g_cells = nullptr, *g_expl = nullptr;
  char* gh_stage = nullptr;  // pinned: GrowIn, then the random numbers
  char* gh_out = nullptr;    // pinned: GrowOut, then the grown chain
  hipGraph_t g_graph = nullptr;
  hipGraphExec_t g_exec = nullptr;
  int cap_k = 0, cap_g = 0, nkb = 0;
  CbmcIn* in = nullptr;
  double *xyz = nullptr, *grown = nullptr, *kpart = nullptr, *out = nullptr, *mz = nullptr;
  double2* cells = nullptr;
  char* h_stage = nullptr;   // pinned: CbmcIn + trials + grown
  double* h_out = nullptr;   // pinned
};

void cbmc_free(CbmcHost* h) {
  if (!h) return;
  if (h->g_exec) (void)hipGraphExecDestroy(h->g_exec);
  if (h->g_graph) (void)hipGraphDestroy(h->g_graph);
  if (h->gh_stage) (void)hipHostFree(h->gh_stage);
  if (h->gh_out) (void)hipHostFree(h->gh_out);
  if (h->h_stage) (void)hipHostFree(h->h_stage);
  if (h->h_out) (void)hipHostFree(h->h_out);
  delete h;
}

namespace {

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
  CHK(dalloc(c, &h->mz, 1));
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
  if (g.dipole && g.use_ewald)
    hipLaunchKernelGGL(k_cbmc_mz, dim3(1), dim3(kBlock), 0, c->st, g, c->d, h->in, h->grown, h->mz);
  hipLaunchKernelGGL(k_cbmc_final, dim3(K), dim3(kBlock), 0, c->st, g, c->d, h->in, h->xyz, h->grown, h->cells, ns,
                     h->kpart, nkb, h->mz, h->out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->h_out, h->out, sizeof(double) * K, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  std::memcpy(energies, h->h_out, sizeof(double) * K);
  return PLUM_OK;
}

namespace {

size_t grow_stage_bytes(int K, int L) { return sizeof(GrowIn) + sizeof(double) * grow_rnd_size(K, L); }
size_t grow_out_bytes(int L) { return sizeof(GrowOut) + sizeof(double) * 6 * (size_t)L; }

int grow_reserve(plum_ctx* c, int K, int L) {
  CbmcHost*& h = c->cbmc;
  if (!h) h = new CbmcHost();
  const int prec = c->p.precision;
  if (K == h->gK && L == h->gL && prec == h->g_prec && h->g_exec) return PLUM_OK;
  if (h->g_exec) (void)hipGraphExecDestroy(h->g_exec);
  if (h->g_graph) (void)hipGraphDestroy(h->g_graph);
  h->g_exec = nullptr;
  h->g_graph = nullptr;
  if (K > h->gK || L > h->gL) {
    // grow-only device buffers (older ones stay with the context until it
    // is destroyed)
    const int k2 = std::max(K, h->gK), l2 = std::max(L, h->gL);
    const int nkb = std::max(1, (c->g.n_kh + kBlock - 1) / kBlock);
    // GrowIn and the random numbers in one buffer, one upload
    char* stage = nullptr;
    CHK(dalloc(c, &stage, grow_stage_bytes(k2, l2)));
    h->g_in = reinterpret_cast<GrowIn*>(stage);
    h->g_rnd = reinterpret_cast<double*>(stage + sizeof(GrowIn));
    CHK(dalloc(c, &h->g_out, 1));
    CHK(dalloc(c, &h->g_grown, 6 * (size_t)l2));
    CHK(dalloc(c, &h->g_Sx, (size_t)std::max(1, c->g.n_kh)));
    CHK(dalloc(c, &h->g_cells, 2 * (size_t)k2 * c->nsplit));
    CHK(dalloc(c, &h->g_expl, (size_t)k2));
    CHK(dalloc(c, &h->g_kpart, (size_t)k2 * nkb));
    if (h->gh_stage) (void)hipHostFree(h->gh_stage);
    if (h->gh_out) (void)hipHostFree(h->gh_out);
    h->gh_stage = h->gh_out = nullptr;
    if (hipHostMalloc((void**)&h->gh_stage, grow_stage_bytes(k2, l2), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->gh_out, grow_out_bytes(l2), hipHostMallocDefault) != hipSuccess)
      return fail(PLUM_E_NOMEM, "hipHostMalloc failed");
  }
  h->gK = K;
  h->gL = L;
  h->g_prec = prec;
  const Geo& g = c->g;
  const int ns = c->nsplit;
  const int nkb = (g.use_ewald && g.n_kh > 0) ? (g.n_kh + kBlock - 1) / kBlock : 0;
  h->g_nkb = nkb;
  const int ncell = 2 * K * ns;
  // one growth: upload, S_x, then two launches per step, download
  HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  (void)hipMemcpyAsync(h->g_in, h->gh_stage, grow_stage_bytes(K, L), hipMemcpyHostToDevice, c->st);
  hipLaunchKernelGGL(k_grow_init, dim3(std::max(1, nkb)), dim3(kBlock), 0, c->st, g, c->d, h->g_in, h->g_Sx,
                     h->g_out);
  for (int s = 0; s < L; s++) {
    if (prec == PLUM_FP32_PAIR)
      hipLaunchKernelGGL(k_grow_trials<1>, dim3(ncell + K + K * nkb), dim3(kBlock), 0, c->st, g, c->d, h->g_in,
                         h->g_rnd, h->g_grown, h->g_Sx, h->g_out, s, ns, h->g_cells, h->g_expl, h->g_kpart);
    else
      hipLaunchKernelGGL(k_grow_trials<0>, dim3(ncell + K + K * nkb), dim3(kBlock), 0, c->st, g, c->d, h->g_in,
                         h->g_rnd, h->g_grown, h->g_Sx, h->g_out, s, ns, h->g_cells, h->g_expl, h->g_kpart);
    hipLaunchKernelGGL(k_grow_select, dim3(1), dim3(kSelThreads), 0, c->st, g, c->d, h->g_in, h->g_rnd,
                       h->g_grown, h->g_Sx, h->g_out, s, ns, nkb, h->g_cells, h->g_expl, h->g_kpart);
  }
  (void)hipMemcpyAsync(h->gh_out, h->g_out, sizeof(GrowOut), hipMemcpyDeviceToHost, c->st);
  (void)hipMemcpyAsync(h->gh_out + sizeof(GrowOut), h->g_grown, sizeof(double) * 6 * (size_t)L,
                       hipMemcpyDeviceToHost, c->st);
  hipGraph_t gr = nullptr;
  HIPCHK(hipStreamEndCapture(c->st, &gr));
  h->g_graph = gr;
  HIPCHK(hipGraphInstantiate(&h->g_exec, h->g_graph, nullptr, nullptr, 0));
  return PLUM_OK;
}

}  // namespace

extern "C" int plum_cbmc_grow(plum_ctx* c, const plum_growth_request* rq, const double* first, const double* disp,
                              const double* ions, const double* sel, const double* actual, double* weight,
                              int* stop_step, double* chain) {
  if (!c || !rq || !weight || !stop_step) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  const int K = rq->n_trials, L = rq->chain_len, ch = rq->charged ? 1 : 0, mode = rq->mode;
  if (mode < 0 || mode > 2) return fail(PLUM_E_ARG, "mode must be 0 (insertion), 1 (deletion) or 2 (test insertion)");
  if (K < 1 || K > kCbmcMaxTrials) return fail(PLUM_E_ARG, "n_trials must be 1..1024");
  if (L < 1 || L > 4096) return fail(PLUM_E_ARG, "chain_len must be 1..4096");
  if (mode == 1 ? !actual : !first) return fail(PLUM_E_ARG, "null first-bead / existing-chain positions");
  if (L > 1 && (!disp || (ch && !ions) || (mode != 1 && !sel))) return fail(PLUM_E_ARG, "null trial draws");
  if (rq->type_chain < 0 || rq->type_chain >= c->g.n_types || (ch && (rq->type_ion < 0 || rq->type_ion >= c->g.n_types)))
    return fail(PLUM_E_ARG, "bead type out of range");
  const bool ex = rq->exclude_lo <= rq->exclude_hi;
  if (ex && (rq->exclude_lo < 0 || rq->exclude_hi >= c->g.n)) return fail(PLUM_E_ARG, "exclusion range out of range");
  CHK(grow_reserve(c, K, L));
  CbmcHost* h = c->cbmc;
  GrowIn in{};
  in.K = K;
  in.L = L;
  in.mode = mode;
  in.charged = ch;
  in.type_c = rq->type_chain;
  in.type_i = ch ? rq->type_ion : rq->type_chain;
  in.ex_lo = ex ? rq->exclude_lo : 0;
  in.ex_hi = ex ? rq->exclude_hi : -1;
  in.q_c = rq->q_chain;
  in.q_i = ch ? rq->q_ion : 0.0;
  std::memcpy(h->gh_stage, &in, sizeof(in));
  double* r = reinterpret_cast<double*>(h->gh_stage + sizeof(GrowIn));
  std::memset(r, 0, sizeof(double) * grow_rnd_size(K, L));
  const size_t lk3 = 3 * (size_t)L * K;
  if (first) std::memcpy(r, first, sizeof(double) * (ch ? 6 : 3));
  if (L > 1) {
    std::memcpy(r + 6 + 3 * (size_t)K, disp + 3 * (size_t)K, sizeof(double) * (lk3 - 3 * (size_t)K));
    if (ch) std::memcpy(r + 6 + lk3 + 3 * (size_t)K, ions + 3 * (size_t)K, sizeof(double) * (lk3 - 3 * (size_t)K));
    if (mode != 1) std::memcpy(r + 6 + 2 * lk3, sel, sizeof(double) * L);
  }
  if (mode == 1) std::memcpy(r + 6 + 2 * lk3 + L, actual, sizeof(double) * 3 * L * (1 + ch));
  for (size_t i = 0; i < grow_rnd_size(K, L); i++)
    if (!std::isfinite(r[i])) return fail(PLUM_E_ARG, "non-finite growth input");
  HIPCHK(hipGraphLaunch(h->g_exec, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  GrowOut o;
  std::memcpy(&o, h->gh_out, sizeof(o));
  *weight = o.W;
  *stop_step = o.stop;
  if (chain) std::memcpy(chain, h->gh_out + sizeof(GrowOut), sizeof(double) * 6 * (size_t)L);
  return PLUM_OK;
}
