// Host side of the device-resident Monte Carlo loop: plum_mc_run (one
// system, run as a batch of one) and plum_batch_* (R systems together),
// include/plum_gpu.h.  Included once, at the end of engine.hip: it shares
// that translation unit's kernels and helpers (build_cells, run_full).
//
// One speculative batch of every system = k_propose_spec, k_plan_spec,
// k_compute_spec, k_reduce_spec, k_resolve_spec (kernels_batch.h).  `batches_per_graph` batches, the cell
// compaction and the status gather are captured once in a HIP graph and
// replayed until every system has reached its move count; the host looks at
// the status between replays (cell overflow: rebuild that system's cell list
// with more room and give it back the skipped moves).

struct plum_batch {
  std::vector<plum_ctx*> reps;
  int R = 0;
  int precision = PLUM_FP64;
  hipStream_t st = nullptr;
  Geo* dG = nullptr;
  Dev* dD = nullptr;
  McDev* dMC = nullptr;
  RepCfg* dRC = nullptr;
  int* d_vb = nullptr;
  int* d_cell_off = nullptr;
  int2* d_status = nullptr;
  int2* h_status = nullptr;   // pinned
  int total_cells = 0;
  int max_nkb_full = 1;
  int n_cu = 0;
  int descr_occ = 1;    // resident k_descr_spec blocks per CU
  int comp_occ = 1;     // resident k_compute_spec blocks per CU
  plum_batch_tuning tune{};
  std::vector<McDev> mc;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_batches = 0, g_grid = 0, g_pmax = 0;  // shape the graph was captured with
  // device pointers the captured kernels received as arguments: the
  // surrogate's BF16 network and the descriptor rows / outputs (they change
  // when the surrogate is enabled again or the row buffers grow)
  const void *g_net = nullptr, *g_sx = nullptr, *g_sf = nullptr;
  bool da = false;        // delayed acceptance (the first system's surrogate, shared)
  // descriptor rows and network outputs of every system's proposals, one
  // slice of 2 kMaxScreen maxm rows per system (R == 1: the surrogate's own)
  float *sx = nullptr, *sf = nullptr;
  int sx_rows = 0, sx_cap = 0, maxm_b = 0;
  int tab_chunks = 1;     // phase-table chunks per proposal (plan kernel rows y > 0)
  int fold_chunks = 0;    // S(k) fold chunks of kPlanThreads k vectors (plan kernel)
  std::vector<int> cell_off_h;   // cell offsets of the systems (prefix, R + 1)
  std::vector<int> grid_gen;     // each system's grid generation at the last upload
  bool profiling = false;
  std::vector<void*> allocs;
};

// Surrogate state of one context (delayed acceptance), defined with its entry
// points in surrogate_api.h.
struct SurHost {
  plum_surrogate_params p{};
  bool on = false;
  NetBF16* net = nullptr;
  float *w32 = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  float *sur_x = nullptr, *sur_f = nullptr;
  float* tr_x = nullptr;
  double* tr_target = nullptr;
  int *tr_n = nullptr, *tr_ctl = nullptr;
  int* h_ctl = nullptr;   // pinned copy of tr_ctl
  double* h_loss = nullptr;  // pinned copy of the last training loss
  int tb_cap = 0;         // minibatch capacity (samples)
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *F = nullptr, *G = nullptr;
  float *D3 = nullptr, *D2 = nullptr, *D1 = nullptr;
  double *T = nullptr, *loss = nullptr;
  int* N = nullptr;
  float* part = nullptr;  // chunk partials of the weight gradients
  int step = 0;
  long long train_steps = 0;
  double last_loss = 0;
  std::vector<void*> allocs;
};

void sur_free(SurHost* s) {
  if (!s) return;
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->h_ctl) (void)hipHostFree(s->h_ctl);
  if (s->h_loss) (void)hipHostFree(s->h_loss);
  delete s;
}

namespace {

int sur_train_step(plum_ctx* c, hipStream_t st, int samples, bool gather, bool update, bool sync);

constexpr int kSpecKernels = 7;
const char* kSpecKernelNames[kSpecKernels] = {"k_propose_spec", "k_descr_spec", "k_mlp_spec", "k_plan_spec",
                                              "k_compute_spec", "k_reduce_spec", "k_resolve_spec"};

template <typename T>
int balloc(plum_batch* b, T** ptr, size_t count) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(1, count));
  if (e != hipSuccess) return fail(PLUM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back(p);
  *ptr = static_cast<T*>(p);
  return PLUM_OK;
}

int batch_grid(const plum_batch* b) { return std::max(1, b->n_cu * b->tune.blocks_per_cu); }

// Proposal slots for batches of up to kMaxScreen screened proposals.
int ensure_screen_slots(plum_ctx* c) {
  if (c->screen_cap >= kMaxScreen) return PLUM_OK;
  Dev& d = c->d;
  const size_t slots = (size_t)kMaxScreen * c->maxm;
  CHK(dalloc(c, &d.mv_old, slots));
  CHK(dalloc(c, &d.mv_new, slots));
  CHK(dalloc(c, &d.mv_type, slots));
  CHK(dalloc(c, &d.smv_idx, slots));
  CHK(dalloc(c, &d.tab, slots * 2 * std::max(1, c->g.tab_w)));
  c->screen_cap = kMaxScreen;
  return PLUM_OK;
}

// Delayed acceptance over a batch: every system screens with the first
// system's network; only the first feeds its training ring.
int batch_setup_da(plum_batch* b) {
  plum_ctx* c0 = b->reps[0];
  b->maxm_b = 0;
  for (int r = 0; r < b->R; r++) {
    CHK(ensure_screen_slots(b->reps[r]));
    b->maxm_b = std::max(b->maxm_b, b->reps[r]->maxm);
  }
  b->sx_rows = 2 * kMaxScreen * b->maxm_b;
  if (b->R == 1) {
    b->sx = c0->sur->sur_x;
    b->sf = c0->sur->sur_f;
    return PLUM_OK;
  }
  const int need = b->R * b->sx_rows;
  if (need > b->sx_cap) {
    CHK(balloc(b, &b->sx, (size_t)need * kDescr));
    CHK(balloc(b, &b->sf, (size_t)need));
    HIPCHK(hipMemsetAsync(b->sx, 0, sizeof(float) * (size_t)need * kDescr, b->st));
    HIPCHK(hipMemsetAsync(b->sf, 0, sizeof(float) * (size_t)need, b->st));
    b->sx_cap = need;
  }
  return PLUM_OK;
}

// Uploads every system's Geo / Dev (a cell rebuild changes both), launch
// shape and Monte Carlo parameters.
void batch_drop_graph(plum_batch* b);

// Cell lists on the tuned grid, and the per-system cell offsets of the
// batch's compaction for the grids the systems have now (a system in two
// batches may have been re-gridded through the other).
int batch_regrid(plum_batch* b) {
  for (plum_ctx* c : b->reps)
    if (b->tune.cell_div > 0 && b->tune.cell_div != c->cell_div) CHK(regrid_cells(c, b->tune.cell_div));
  std::vector<int> off(b->R + 1, 0);
  for (int r = 0; r < b->R; r++) off[r + 1] = off[r] + b->reps[r]->g.ncell;
  if (off == b->cell_off_h) return PLUM_OK;
  b->cell_off_h = off;
  b->total_cells = off[b->R];
  HIPCHK(hipMemcpy(b->d_cell_off, off.data(), sizeof(int) * (b->R + 1), hipMemcpyHostToDevice));
  batch_drop_graph(b);   // the compaction grid changed
  return PLUM_OK;
}

// True when a system's cell grid changed since the batch uploaded its state.
bool batch_grid_moved(const plum_batch* b) {
  for (int r = 0; r < b->R; r++)
    if (b->grid_gen[r] != b->reps[r]->grid_gen) return true;
  return false;
}

int batch_upload_state(plum_batch* b) {
  CHK(batch_regrid(b));
  b->grid_gen.assign(b->R, 0);
  for (int r = 0; r < b->R; r++) b->grid_gen[r] = b->reps[r]->grid_gen;
  std::vector<Geo> G(b->R);
  std::vector<Dev> D(b->R);
  std::vector<RepCfg> RCv(b->R);
  b->max_nkb_full = 1;
  b->da = b->reps[0]->sur && b->reps[0]->sur->on;
  if (b->da) CHK(batch_setup_da(b));
  for (int r = 0; r < b->R; r++) {
    plum_ctx* c = b->reps[r];
    c->nsplit = split_for(c->cand, b->tune.split_div);
    G[r] = c->g;
    D[r] = c->d;
    RepCfg& q = RCv[r];
    q.nkb_s = c->nkb_s;
    q.nrun = c->nrun;
    q.qfill = b->n_cu * std::min(b->comp_occ, b->tune.blocks_per_cu) * kQWaves;
    q.nsplit = c->nsplit;
    q.ncell = c->g.ncell;
    q.pmax = b->tune.pmax;
    q.maxm = c->maxm;
    q.nmmb = c->nmmb;
    q.rstride = 2 * c->maxm * kMaxQWave;
    q.xcap = c->xcap;
    q.da = b->da ? 1 : 0;
    q.train = (b->da && r == 0) ? 1 : 0;
    q.fill = q.train ? b->reps[0]->sur->p.fill : 0;
    q.r_d = q.da ? (float)b->reps[0]->sur->p.r_d : 0.f;
    if (b->da) {
      D[r].sur_x = b->sx + (size_t)r * b->sx_rows * kDescr;
      D[r].sur_f = b->sf + (size_t)r * b->sx_rows;
    }
    b->max_nkb_full = std::max(b->max_nkb_full, c->nkb_full);
  }
  HIPCHK(hipMemcpyAsync(b->dG, G.data(), sizeof(Geo) * b->R, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->dD, D.data(), sizeof(Dev) * b->R, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->dRC, RCv.data(), sizeof(RepCfg) * b->R, hipMemcpyHostToDevice, b->st));
  if (!b->mc.empty())
    HIPCHK(hipMemcpyAsync(b->dMC, b->mc.data(), sizeof(McDev) * b->R, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  return PLUM_OK;
}

// One speculative batch of every system; with `ev` (profiling) each kernel is
// bracketed by events ev[k], ev[k + 1].
void launch_batch(plum_batch* b, hipEvent_t* ev = nullptr) {
  const int R = b->R;
  if (ev) (void)hipEventRecord(ev[0], b->st);
  hipLaunchKernelGGL(k_propose_spec, dim3(b->tune.pmax, R), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dMC, b->dRC,
                     R);
  if (ev) (void)hipEventRecord(ev[1], b->st);
  if (b->da) {
    // surrogate: descriptors of every moved bead (old, new) of every system,
    // then the network on all of them in one launch
    plum_ctx* c = b->reps[0];
    // descriptor rows: four waves per row for a lone system (few rows, a
    // long candidate scan each), one per row when many systems fill the chip;
    // as many blocks over all systems as the chip holds at once (a second
    // round of blocks would start only as the first finishes its rows),
    // never more than the rows
    const int wpr = R == 1 ? 4 : (R < 8 ? 2 : 1);
    const int rows_max = 2 * b->tune.pmax * (R == 1 ? c->maxm : b->maxm_b);
    const int per_block = kDescrWaves / wpr;
    const int dblocks =
        std::max(1, std::min((rows_max + per_block - 1) / per_block, b->descr_occ * b->n_cu / R));
    hipLaunchKernelGGL(k_descr_spec, dim3(dblocks * R), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dRC, wpr, R,
                       dblocks);
    if (ev) (void)hipEventRecord(ev[2], b->st);
    // the network on the moved beads' rows only: tiles of 32 rows, at most
    // 2 pmax maxm rows per system, blocks past a system's tiles exit
    const int tiles = (rows_max + 31) / 32;
    const int mblocks = std::max(1, std::min(tiles, (b->n_cu + R - 1) / R));
    hipLaunchKernelGGL(k_mlp_spec, dim3(mblocks, R), dim3(kBlock), 0, b->st, c->sur->net, b->sx, b->sf, b->dD,
                       b->dRC, b->sx_rows);
  } else if (ev) {
    (void)hipEventRecord(ev[2], b->st);
  }
  if (ev) (void)hipEventRecord(ev[3], b->st);
  // y = 0: the plan of system x; then chunks of its phase tables; then chunks
  // of its S(k) fold
  const int ntab = b->tune.pmax * b->tab_chunks;
  hipLaunchKernelGGL(k_plan_spec, dim3(R, 1 + ntab + b->fold_chunks), dim3(kPlanThreads), 0, b->st, b->dG, b->dD,
                     b->dMC, b->dRC, b->d_vb, R, b->tab_chunks, ntab);
  if (ev) (void)hipEventRecord(ev[4], b->st);
  if (b->precision == PLUM_FP32_PAIR)
    hipLaunchKernelGGL(k_compute_spec<1>, dim3(batch_grid(b)), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dRC,
                       b->d_vb, R);
  else
    hipLaunchKernelGGL(k_compute_spec<0>, dim3(batch_grid(b)), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dRC,
                       b->d_vb, R);
  if (ev) (void)hipEventRecord(ev[5], b->st);
  hipLaunchKernelGGL(k_reduce_spec, dim3(kKVals + kMaxProp + kMaxPairs, R), dim3(kReduceThreads), 0, b->st, b->dG,
                     b->dD, b->dRC, R);
  if (ev) (void)hipEventRecord(ev[6], b->st);
  hipLaunchKernelGGL(k_resolve_spec, dim3(R), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dMC, b->dRC, R);
  if (ev) (void)hipEventRecord(ev[7], b->st);
}

void launch_tail(plum_batch* b) {
  hipLaunchKernelGGL(k_compact_batch, dim3((b->total_cells + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dG,
                     b->dD, b->d_cell_off, b->R);
  hipLaunchKernelGGL(k_batch_status, dim3((b->R + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dD, b->dMC, b->R,
                     b->d_status);
}

void batch_drop_graph(plum_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->g_batches = b->g_grid = b->g_pmax = 0;
  b->g_net = b->g_sx = b->g_sf = nullptr;
}

// The surrogate network a batch's launches read (null without delayed
// acceptance).
const void* batch_net(const plum_batch* b) {
  return b->da && b->reps[0]->sur ? (const void*)b->reps[0]->sur->net : nullptr;
}

// True when the captured graph no longer matches what launch_batch would
// launch now: its shape or any device pointer baked into its arguments.
bool graph_stale(const plum_batch* b, int nb) {
  const int key_pmax = b->tune.pmax + (b->da ? 1000 : 0);
  return !b->gexec || b->g_batches != nb || b->g_grid != batch_grid(b) || b->g_pmax != key_pmax ||
         b->g_net != batch_net(b) || b->g_sx != (b->da ? (const void*)b->sx : nullptr) ||
         b->g_sf != (b->da ? (const void*)b->sf : nullptr);
}

// `nb` batches of every system, then compaction and the status gather.  With
// profiling, direct launches bracketed by events (per-kernel device times).
int batch_replay(plum_batch* b, int nb, bool use_graph, plum_ctx* prof) {
  if (prof) {
    std::vector<hipEvent_t> ev(kSpecKernels + 1);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    for (int i = 0; i < nb; i++) {
      launch_batch(b, ev.data());
      HIPCHK(hipEventSynchronize(ev[kSpecKernels]));
      for (int k = 0; k < kSpecKernels; k++) {
        if (!b->da && (k == 1 || k == 2)) continue;   // no surrogate kernels
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) {
          prof->ksum_us[k] += 1000.0 * ms;
          prof->kcount[k] += 1;
        }
      }
      // stop once every system has made its moves (no idle batches timed)
      hipLaunchKernelGGL(k_batch_status, dim3((b->R + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dD, b->dMC,
                         b->R, b->d_status);
      HIPCHK(hipMemcpyAsync(b->h_status, b->d_status, sizeof(int2) * b->R, hipMemcpyDeviceToHost, b->st));
      HIPCHK(hipStreamSynchronize(b->st));
      bool busy = false;
      for (int r = 0; r < b->R; r++) busy |= b->h_status[r].y > 0 && !b->h_status[r].x;
      if (!busy) break;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    launch_tail(b);
    HIPCHK(hipGetLastError());
    return PLUM_OK;
  }
  if (!use_graph) {
    for (int i = 0; i < nb; i++) launch_batch(b);
    launch_tail(b);
    HIPCHK(hipGetLastError());
    return PLUM_OK;
  }
  if (graph_stale(b, nb)) {
    batch_drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < nb; i++) launch_batch(b);
    launch_tail(b);
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamEndCapture(b->st, &gr));
    b->graph = gr;
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->g_batches = nb;
    b->g_grid = batch_grid(b);
    b->g_pmax = b->tune.pmax + (b->da ? 1000 : 0);
    b->g_net = batch_net(b);
    b->g_sx = b->da ? b->sx : nullptr;
    b->g_sf = b->da ? b->sf : nullptr;
  }
  HIPCHK(hipGraphLaunch(b->gexec, b->st));
  return PLUM_OK;
}

// Pending S(k) of every system folded in (end of a run, before a rebuild).
int batch_flush(plum_batch* b) {
  hipLaunchKernelGGL(k_flush_spec, dim3(std::min(b->max_nkb_full, 64), b->R), dim3(kBlock), 0, b->st, b->dG, b->dD,
                     b->R);
  hipLaunchKernelGGL(k_clear_mask, dim3((b->R + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dD, b->R);
  HIPCHK(hipGetLastError());
  return PLUM_OK;
}

int read_ctl(plum_ctx* c, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  return PLUM_OK;
}

// Chain / ion lists (simulation.cc:247-266: from the phantom index on) and
// the McDev of one system.
int mc_setup(plum_ctx* c, const plum_mc_params* mcp, hipStream_t st, McDev* out) {
  double tot = 0;
  for (int k = 0; k < 5; k++) {
    if (!(mcp->move_prob[k] >= 0)) return fail(PLUM_E_ARG, "negative move probability");
    tot += mcp->move_prob[k];
  }
  if (std::fabs(tot - 1.0) > 1e-4) return fail(PLUM_E_ARG, "move probabilities must sum to 1 (simulation.cc:158-166)");
  if (mcp->phantom < 0 || mcp->phantom > c->g.n_mols) return fail(PLUM_E_ARG, "phantom count out of range");
  std::vector<int> chains, ions;
  for (int m = mcp->phantom; m < c->g.n_mols; m++) {
    if (c->mol_start[m + 1] - c->mol_start[m] > 1) chains.push_back(m);
    else ions.push_back(m);
  }
  if (!chains.empty())
    HIPCHK(hipMemcpyAsync(c->d.chain_ids, chains.data(), sizeof(int) * chains.size(), hipMemcpyHostToDevice, st));
  if (!ions.empty())
    HIPCHK(hipMemcpyAsync(c->d.ion_ids, ions.data(), sizeof(int) * ions.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  McDev mc{};
  for (int k = 0; k < 5; k++) mc.cum_prob[k] = mcp->move_prob[k];
  mc.move_size = mcp->move_size;
  mc.bond_len = mcp->bond_len;
  mc.vary_bond = mcp->vary_bond;
  mc.n_chains = (int)chains.size();
  mc.n_ions = (int)ions.size();
  mc.seed = mcp->seed;
  mc.ctr_limit = 0;
  *out = mc;
  return PLUM_OK;
}

int batch_new(plum_ctx* const* replicas, int n, plum_batch** out) {
  plum_batch* b = new plum_batch();
  b->reps.assign(replicas, replicas + n);
  b->R = n;
  b->precision = replicas[0]->p.precision;
  b->tune.blocks_per_cu = 8;
  b->tune.split_div = kDefaultSplitDiv;
  b->tune.batches_per_graph = 32;
  b->tune.pmax = kMaxProp;
  b->tune.sweep_us = 0;
  b->tune.cell_div = 0;
  int dev = 0;
  int rc = PLUM_OK;
  auto bail = [&](int code) {
    plum_batch_destroy(b);
    return code;
  };
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "cannot query the device"));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->descr_occ, k_descr_spec, kBlock, 0) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "cannot query the occupancy of the descriptor kernel"));
  b->descr_occ = std::max(1, b->descr_occ);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->comp_occ,
                                                   b->precision == PLUM_FP32_PAIR ? k_compute_spec<1> : k_compute_spec<0>,
                                                   kBlock, 0) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "cannot query the occupancy of the compute kernel"));
  b->comp_occ = std::max(1, b->comp_occ);
  if (hipStreamCreate(&b->st) != hipSuccess) return bail(fail(PLUM_E_HIP, "hipStreamCreate failed"));
  if ((rc = balloc(b, &b->dG, n)) || (rc = balloc(b, &b->dD, n)) || (rc = balloc(b, &b->dMC, n)) ||
      (rc = balloc(b, &b->dRC, n)) || (rc = balloc(b, &b->d_vb, n)) || (rc = balloc(b, &b->d_cell_off, n + 1)) ||
      (rc = balloc(b, &b->d_status, n)))
    return bail(rc);
  if (hipHostMalloc((void**)&b->h_status, sizeof(int2) * n, hipHostMallocDefault) != hipSuccess)
    return bail(fail(PLUM_E_NOMEM, "hipHostMalloc failed"));
  std::vector<int> off(n + 1, 0);
  for (int r = 0; r < n; r++) off[r + 1] = off[r] + replicas[r]->g.ncell;
  b->total_cells = off[n];
  b->cell_off_h = off;
  for (int r = 0; r < n; r++) {
    const int cr = tab_chunk_rows(replicas[r]->g.tab_w);
    b->tab_chunks = std::max(b->tab_chunks, (2 * replicas[r]->maxm + cr - 1) / cr);
    b->fold_chunks = std::max(b->fold_chunks, (replicas[r]->g.n_kh + kPlanThreads - 1) / kPlanThreads);
  }
  if (hipMemcpy(b->d_cell_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "upload of the cell offsets failed"));
  for (plum_ctx* c : b->reps)
    if (hipStreamSynchronize(c->st) != hipSuccess) return bail(fail(PLUM_E_HIP, "replica stream failed"));
  if ((rc = batch_upload_state(b)) != PLUM_OK) return bail(rc);
  *out = b;
  return PLUM_OK;
}

// The device loop proper: n_moves trial moves of every system.
int batch_run(plum_batch* b, const plum_mc_params* mcp, long long n_moves, bool use_graph, plum_ctx* prof,
              plum_batch_stats* total, plum_mc_stats* per) {
  const int R = b->R;
  for (int r = 0; r < R; r++)
    if (b->reps[r]->pending) return fail(PLUM_E_STATE, "a system has a pending trial move");
  // work still queued on a system's own stream (a rollback's commit kernel,
  // a host-path upload) comes before this batch's kernels on the batch stream
  for (int r = 0; r < R; r++) {
    plum_ctx* c = b->reps[r];
    HIPCHK(hipEventRecord(c->ev_idle, c->st));
    HIPCHK(hipStreamWaitEvent(b->st, c->ev_idle, 0));
  }
  if (batch_grid_moved(b)) CHK(batch_upload_state(b));
  b->mc.assign(R, McDev{});
  for (int r = 0; r < R; r++) CHK(mc_setup(b->reps[r], &mcp[r], b->st, &b->mc[r]));
  std::vector<unsigned long long> m0(R);
  std::vector<long long> sk0(R), att0(5 * (size_t)R), acc0(5 * (size_t)R);
  for (int r = 0; r < R; r++) CHK(read_ctl(b->reps[r], b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  for (int r = 0; r < R; r++) {
    const Ctl* h = b->reps[r]->h_ctl;
    m0[r] = h->move_counter;
    sk0[r] = h->skipped;
    for (int k = 0; k < 5; k++) {
      att0[5 * (size_t)r + k] = h->attempted[k];
      acc0[5 * (size_t)r + k] = h->accepted[k];
    }
    b->mc[r].ctr_limit = m0[r] + (unsigned long long)n_moves;
  }
  CHK(batch_upload_state(b));
  long long resync_every = 0;
  for (int r = 0; r < R; r++)
    if (mcp[r].resync_every > 0)
      resync_every = resync_every ? std::min<long long>(resync_every, mcp[r].resync_every) : mcp[r].resync_every;
  std::vector<long long> credited(R, 0);   // skipped moves already given back
  const auto t0 = std::chrono::steady_clock::now();
  long long since_resync = 0, rebuilds = 0;
  for (;;) {
    CHK(batch_replay(b, b->tune.batches_per_graph, use_graph, prof));
    HIPCHK(hipMemcpyAsync(b->h_status, b->d_status, sizeof(int2) * R, hipMemcpyDeviceToHost, b->st));
    if (b->da) HIPCHK(hipMemcpyAsync(b->reps[0]->sur->h_ctl, b->reps[0]->sur->tr_ctl, 2 * sizeof(int),
                                     hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    if (b->da) {
      // online training on the newest exact samples
      SurHost* su = b->reps[0]->sur;
      if (su->p.train_every > 0 && su->h_ctl[0] >= su->p.train_samples)
        CHK(sur_train_step(b->reps[0], b->st, su->p.train_samples, true, true, false));
    }
    since_resync += (long long)b->tune.batches_per_graph * b->tune.pmax;
    bool busy = false, fix = false;
    for (int r = 0; r < R; r++) {
      busy |= b->h_status[r].y > 0;
      fix |= b->h_status[r].x != 0;
    }
    const bool resync = resync_every > 0 && since_resync >= resync_every && busy;
    if (fix || resync) {
      CHK(batch_flush(b));
      HIPCHK(hipStreamSynchronize(b->st));
      for (int r = 0; r < R; r++) {
        plum_ctx* c = b->reps[r];
        if (b->h_status[r].x) {
          // cell overflow: more room, and the moves skipped since get redone
          if (++rebuilds > 16L * R) return fail(PLUM_E_CAPACITY, "cell list keeps overflowing");
          int zero = 0;
          HIPCHK(hipMemcpyAsync(&c->d.ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, b->st));
          CHK(read_ctl(c, b->st));
          HIPCHK(hipStreamSynchronize(b->st));
          CHK(build_cells(c, 2 * c->g.cap));
          const long long sk = c->h_ctl->skipped - sk0[r];
          b->mc[r].ctr_limit += (unsigned long long)(sk - credited[r]);
          credited[r] = sk;
          busy = true;
        } else if (resync) {
          CHK(build_cells(c, c->g.cap));
          double f[8];
          CHK(run_full(c, 1, f));
        }
      }
      if (resync) since_resync = 0;
      CHK(batch_upload_state(b));
    }
    if (!busy) break;
  }
  CHK(batch_flush(b));
  hipLaunchKernelGGL(k_compact_batch, dim3((b->total_cells + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dG,
                     b->dD, b->d_cell_off, b->R);
  HIPCHK(hipGetLastError());
  for (int r = 0; r < R; r++) CHK(read_ctl(b->reps[r], b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  const auto t1 = std::chrono::steady_clock::now();
  const double secs = std::chrono::duration<double>(t1 - t0).count();
  plum_batch_stats st{};
  for (int r = 0; r < R; r++) {
    const Ctl* h = b->reps[r]->h_ctl;
    const long long skipped = h->skipped - sk0[r];
    const long long moves = ((long long)h->move_counter - (long long)m0[r]) - skipped;
    st.moves += moves;
    for (int k = 0; k < 5; k++) {
      st.attempted[k] += h->attempted[k] - att0[5 * (size_t)r + k];
      st.accepted[k] += h->accepted[k] - acc0[5 * (size_t)r + k];
    }
    if (skipped > 0) st.overflow_replicas += 1;
    if (per) {
      plum_mc_stats& p = per[r];
      std::memset(&p, 0, sizeof(p));
      for (int k = 0; k < 5; k++) {
        p.attempted[k] = h->attempted[k] - att0[5 * (size_t)r + k];
        p.accepted[k] = h->accepted[k] - acc0[5 * (size_t)r + k];
      }
      p.moves = moves;
      p.seconds = secs;
      to_terms(h->running, &p.totals);
    }
  }
  st.sweeps = n_moves;
  st.seconds = secs;
  if (total) *total = st;
  return PLUM_OK;
}

int check_tuning(const plum_batch_tuning* t) {
  if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return fail(PLUM_E_ARG, "blocks_per_cu must be 1..64");
  if (t->split_div < 16) return fail(PLUM_E_ARG, "split_div must be >= 16");
  if (t->batches_per_graph < 1 || t->batches_per_graph > 1024)
    return fail(PLUM_E_ARG, "batches_per_graph must be 1..1024");
  if (t->pmax < 1 || t->pmax > kMaxScreen) return fail(PLUM_E_ARG, "pmax must be 1..32");
  if (t->cell_div < 0 || t->cell_div > 8) return fail(PLUM_E_ARG, "cell_div must be 0..8");
  return PLUM_OK;
}

// Proposals per batch: at most the exact slots without the surrogate, at
// most the allocated proposal slots with it.
int check_pmax(const plum_batch* b, int pmax) {
  for (plum_ctx* c : b->reps) {
    const bool da = b->reps[0]->sur && b->reps[0]->sur->on;
    if (pmax > (da ? c->screen_cap : kMaxProp))
      return fail(PLUM_E_ARG, da ? "pmax exceeds the screened proposals of the surrogate" : "pmax must be 1..16");
  }
  return PLUM_OK;
}

}  // namespace

extern "C" {

int plum_mc_run(plum_ctx* c, const plum_mc_params* mcp, long long n_moves, plum_mc_stats* stats) {
  if (!c || !mcp) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  if (n_moves < 0) return fail(PLUM_E_ARG, "negative move count");
  if (!c->self) {
    plum_ctx* one[1] = {c};
    CHK(batch_new(one, 1, &c->self));
  }
  if (c->profiling) {
    c->kname.assign(kSpecKernelNames, kSpecKernelNames + kSpecKernels);
    c->ksum_us.assign(kSpecKernels, 0.0);
    c->kcount.assign(kSpecKernels, 0);
  }
  plum_mc_stats st{};
  CHK(batch_run(c->self, mcp, n_moves, true, c->profiling ? c : nullptr, nullptr, &st));
  if (stats) *stats = st;
  return PLUM_OK;
}

int plum_mc_record(plum_ctx* c, int capacity) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (capacity < 0 || capacity > (1 << 24)) return fail(PLUM_E_ARG, "record capacity must be 0..2^24");
  if (capacity == 0) {
    c->d.drec = nullptr;
    c->d.drec_pos = nullptr;
    c->d.drec_idx = nullptr;
    c->d.drec_ctl = nullptr;
    return PLUM_OK;
  }
  HIPCHK(hipStreamSynchronize(c->st));
  if (c->self) HIPCHK(hipStreamSynchronize(c->self->st));
  if (capacity > c->drec_cap || !c->d.drec_ctl) {
    Dev& d = c->d;
    CHK(dalloc(c, &d.drec, (size_t)capacity));
    CHK(dalloc(c, &d.drec_pos, (size_t)capacity * c->maxm));
    CHK(dalloc(c, &d.drec_idx, (size_t)capacity * c->maxm));
    CHK(dalloc(c, &d.drec_ctl, 2));
    c->drec_cap = capacity;
  }
  const int ctl[2] = {0, capacity};
  HIPCHK(hipMemcpyAsync(c->d.drec_ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return PLUM_OK;
}

int plum_mc_get_records(plum_ctx* c, int cap, plum_decision* out, int* idx, double* xyz, long long* n_decided) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (!c->d.drec_ctl) return fail(PLUM_E_STATE, "decision record not armed (plum_mc_record)");
  if (cap < 0 || (cap > 0 && (!out || !idx || !xyz))) return fail(PLUM_E_ARG, "null output");
  int ctl[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(ctl, c->d.drec_ctl, sizeof(ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (n_decided) *n_decided = ctl[0];
  const int n = std::min(std::min(ctl[0], ctl[1]), cap);
  if (n <= 0) return 0;
  const int M = c->maxm;
  std::vector<DecRec> r(n);
  std::vector<double4> pos((size_t)n * M);
  std::vector<int> id((size_t)n * M);
  HIPCHK(hipMemcpyAsync(r.data(), c->d.drec, sizeof(DecRec) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(pos.data(), c->d.drec_pos, sizeof(double4) * pos.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(id.data(), c->d.drec_idx, sizeof(int) * id.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int k = 0; k < n; k++) {
    plum_decision& o = out[k];
    const DecRec& s = r[k];
    o.move = s.move;
    o.mol = s.mol;
    o.n_moved = s.n;
    o.move_type = s.type;
    o.accepted = s.accepted;
    o.evaluated = s.evaluated;
    o.early = s.early;
    o.first_stage_pass = s.pass;
    o.lo = s.lo;
    o.hi = s.hi;
    to_terms(s.terms, &o.dE);
    o.total = s.total;
    o.dE_surrogate = s.de_s;
    for (int m = 0; m < M; m++) {
      const size_t j = (size_t)k * M + m;
      const bool in = m < s.n;
      idx[j] = in ? id[j] : -1;
      xyz[3 * j] = in ? pos[j].x : 0.0;
      xyz[3 * j + 1] = in ? pos[j].y : 0.0;
      xyz[3 * j + 2] = in ? pos[j].z : 0.0;
    }
  }
  return n;
}

int plum_mc_get_tuning(plum_ctx* c, plum_batch_tuning* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  if (!c->self) {
    plum_ctx* one[1] = {c};
    CHK(batch_new(one, 1, &c->self));
  }
  *out = c->self->tune;
  return PLUM_OK;
}

int plum_mc_set_tuning(plum_ctx* c, const plum_batch_tuning* t) {
  if (!c || !t) return fail(PLUM_E_ARG, "null argument");
  CHK(check_tuning(t));
  if (!c->self) {
    plum_ctx* one[1] = {c};
    CHK(batch_new(one, 1, &c->self));
  }
  CHK(check_pmax(c->self, t->pmax));
  const double us = c->self->tune.sweep_us;
  c->self->tune = *t;
  c->self->tune.sweep_us = us;
  CHK(batch_upload_state(c->self));
  return PLUM_OK;
}

int plum_batch_create(plum_ctx* const* replicas, int n, plum_batch** out) {
  if (!out || !replicas) return fail(PLUM_E_ARG, "null argument");
  *out = nullptr;
  if (n <= 0 || n > kMaxReplicas) return fail(PLUM_E_ARG, "replica count must be 1..1024");
  for (int r = 0; r < n; r++) {
    if (!replicas[r]) return fail(PLUM_E_ARG, "null replica");
    if (replicas[r]->pending) return fail(PLUM_E_STATE, "a replica has a pending trial move");
    if (replicas[r]->p.precision != replicas[0]->p.precision)
      return fail(PLUM_E_ARG, "all replicas must use the same precision");
    for (int q = 0; q < r; q++)
      if (replicas[q] == replicas[r]) return fail(PLUM_E_ARG, "replica listed twice");
  }
  return batch_new(replicas, n, out);
}

int plum_batch_destroy(plum_batch* b) {
  if (!b) return PLUM_OK;
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_drop_graph(b);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->h_status) (void)hipHostFree(b->h_status);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
  return PLUM_OK;
}

int plum_batch_get_tuning(const plum_batch* b, plum_batch_tuning* out) {
  if (!b || !out) return fail(PLUM_E_ARG, "null argument");
  *out = b->tune;
  return PLUM_OK;
}

int plum_batch_set_tuning(plum_batch* b, const plum_batch_tuning* t) {
  if (!b || !t) return fail(PLUM_E_ARG, "null argument");
  CHK(check_tuning(t));
  CHK(check_pmax(b, t->pmax));
  const double us = b->tune.sweep_us;
  b->tune = *t;
  b->tune.sweep_us = us;
  CHK(batch_upload_state(b));
  return PLUM_OK;
}

int plum_batch_mc_run(plum_batch* b, const plum_mc_params* mcp, long long n_moves, int use_graph,
                      plum_batch_stats* total, plum_mc_stats* per) {
  if (!b || !mcp) return fail(PLUM_E_ARG, "null argument");
  if (n_moves < 0) return fail(PLUM_E_ARG, "negative move count");
  return batch_run(b, mcp, n_moves, use_graph != 0, nullptr, total, per);
}

// Times `moves_per_trial` moves of every system for each candidate launch
// shape (work-groups per CU of k_compute_spec, candidate slots per
// real-space query block, proposals per speculative batch) and keeps the
// fastest.  The trial moves are ordinary Monte Carlo moves.
int plum_batch_tune(plum_batch* b, const plum_mc_params* mcp, int moves_per_trial, plum_batch_tuning* best) {
  if (!b || !mcp) return fail(PLUM_E_ARG, "null argument");
  if (moves_per_trial < 1) return fail(PLUM_E_ARG, "moves_per_trial must be >= 1");
  const int cdiv[] = {1, 2, 3};
  const int bpc[] = {4, 8, 16};
  const int div[] = {192, 384, 768};
  const int pm[] = {1, 2, 4, 8, 16};
  constexpr int kRepeats = 3;
  const plum_batch_tuning keep = b->tune;
  plum_batch_tuning win = keep;
  double best_s = 1e300;
  for (int c : cdiv)
    for (int i : bpc)
      for (int j : div)
        for (int k : pm) {
          plum_batch_tuning t = keep;
          t.cell_div = c;
          t.blocks_per_cu = i;
          t.split_div = j;
          t.pmax = k;
          t.batches_per_graph = std::max(1, std::min(64, moves_per_trial / k));
          CHK(plum_batch_set_tuning(b, &t));
          // one untimed pass (graph capture), then the median of three
          plum_batch_stats st{};
          CHK(batch_run(b, mcp, moves_per_trial, true, nullptr, &st, nullptr));
          double s3[kRepeats];
          for (int rep = 0; rep < kRepeats; rep++) {
            st = plum_batch_stats{};
            CHK(batch_run(b, mcp, moves_per_trial, true, nullptr, &st, nullptr));
            s3[rep] = st.seconds;
          }
          std::sort(s3, s3 + kRepeats);
          const double med = s3[kRepeats / 2];
          if (med < best_s) {
            best_s = med;
            win = t;
            win.sweep_us = 1e6 * med / moves_per_trial;
          }
        }
  win.batches_per_graph = keep.batches_per_graph;
  CHK(plum_batch_set_tuning(b, &win));
  b->tune.sweep_us = win.sweep_us;
  if (best) *best = b->tune;
  return PLUM_OK;
}

}  // extern "C"
