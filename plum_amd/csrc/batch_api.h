// Host side of the batched-replica entry points (plum_batch_* in
// include/plum_gpu.h).  Included once, at the end of engine.hip: it shares
// that translation unit's kernels and helpers (build_cells, flush_sk, the
// control-block copies).
//
// A sweep = one trial move of every replica: k_compute_batch (all replicas'
// k-space / real-space / moved-moved blocks over one fixed grid) followed by
// k_serial_batch (one block per replica).  `sweeps_per_graph` sweeps, the
// cell compaction and the overflow-flag gather are captured once in a HIP
// graph and replayed; the host looks at the flags between replays.

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
  int* d_flags = nullptr;
  int* h_flags = nullptr;     // pinned
  int total_cells = 0;
  int n_cu = 0;
  plum_batch_tuning tune{};
  std::vector<RepCfg> rc;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_sweeps = 0, g_grid = 0;  // shape the graph was captured with
  std::vector<void*> allocs;
};

namespace {

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

// Uploads every replica's Geo / Dev (a cell rebuild can change both) and the
// launch shapes.
int batch_upload_state(plum_batch* b) {
  std::vector<Geo> G(b->R);
  std::vector<Dev> D(b->R);
  b->rc.resize(b->R);
  for (int r = 0; r < b->R; r++) {
    plum_ctx* c = b->reps[r];
    c->nsplit = split_for(c->cand, b->tune.split_div);
    G[r] = c->g;
    D[r] = c->d;
    RepCfg& q = b->rc[r];
    q.nkb = (c->g.use_ewald && c->g.n_kh > 0) ? c->nkb : 0;
    q.nsplit = c->nsplit;
    q.ncell = c->g.ncell;
    q.pad = 0;
  }
  HIPCHK(hipMemcpyAsync(b->dG, G.data(), sizeof(Geo) * b->R, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->dD, D.data(), sizeof(Dev) * b->R, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->dRC, b->rc.data(), sizeof(RepCfg) * b->R, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  return PLUM_OK;
}

// `sweeps` trial moves of every replica, then compaction and the flag gather.
void batch_enqueue(plum_batch* b, int sweeps) {
  const int R = b->R, grid = batch_grid(b);
  hipLaunchKernelGGL(k_serial_batch, dim3(R), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dMC, b->dRC, b->d_vb, R,
                     0, 1);
  for (int i = 0; i < sweeps; i++) {
    if (b->precision == PLUM_FP32_PAIR)
      hipLaunchKernelGGL(k_compute_batch<1>, dim3(grid), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dRC, b->d_vb, R);
    else
      hipLaunchKernelGGL(k_compute_batch<0>, dim3(grid), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dRC, b->d_vb, R);
    hipLaunchKernelGGL(k_serial_batch, dim3(R), dim3(kBlock), 0, b->st, b->dG, b->dD, b->dMC, b->dRC, b->d_vb, R,
                       1, i < sweeps - 1 ? 1 : 0);
  }
  hipLaunchKernelGGL(k_compact_batch, dim3((b->total_cells + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dG,
                     b->dD, b->d_cell_off, R);
  hipLaunchKernelGGL(k_batch_flags, dim3((R + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dD, R, b->d_flags);
}

void batch_drop_graph(plum_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->g_sweeps = 0;
  b->g_grid = 0;
}

int batch_launch(plum_batch* b, int sweeps, bool use_graph) {
  if (!use_graph) {
    batch_enqueue(b, sweeps);
    HIPCHK(hipGetLastError());
    return PLUM_OK;
  }
  if (!b->gexec || b->g_sweeps != sweeps || b->g_grid != batch_grid(b)) {
    batch_drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    batch_enqueue(b, sweeps);
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamEndCapture(b->st, &gr));
    b->graph = gr;
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->g_sweeps = sweeps;
    b->g_grid = batch_grid(b);
  }
  HIPCHK(hipGraphLaunch(b->gexec, b->st));
  return PLUM_OK;
}

int batch_flush(plum_batch* b) {
  int maxk = 1;
  for (plum_ctx* c : b->reps) maxk = std::max(maxk, c->nkb_full);
  hipLaunchKernelGGL(k_flush_sk_batch, dim3(std::min(maxk, 64), b->R), dim3(kBlock), 0, b->st, b->dG, b->dD, b->R);
  hipLaunchKernelGGL(k_clear_sk_pending_batch, dim3((b->R + kBlock - 1) / kBlock), dim3(kBlock), 0, b->st, b->dD,
                     b->R);
  HIPCHK(hipGetLastError());
  return PLUM_OK;
}

int read_ctl(plum_ctx* c, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  return PLUM_OK;
}

// Chain / ion lists and the McDev of one replica (as plum_mc_run sets them).
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
  *out = mc;
  return PLUM_OK;
}

}  // namespace

extern "C" {

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
  plum_batch* b = new plum_batch();
  b->reps.assign(replicas, replicas + n);
  b->R = n;
  b->precision = replicas[0]->p.precision;
  b->tune.blocks_per_cu = 8;
  b->tune.split_div = kDefaultSplitDiv;
  b->tune.sweeps_per_graph = 256;
  b->tune.sweep_us = 0;
  int dev = 0;
  int rc = PLUM_OK;
  auto bail = [&](int code) {
    plum_batch_destroy(b);
    return code;
  };
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "cannot query the device"));
  if (hipStreamCreate(&b->st) != hipSuccess) return bail(fail(PLUM_E_HIP, "hipStreamCreate failed"));
  if ((rc = balloc(b, &b->dG, n)) || (rc = balloc(b, &b->dD, n)) || (rc = balloc(b, &b->dMC, n)) ||
      (rc = balloc(b, &b->dRC, n)) || (rc = balloc(b, &b->d_vb, n)) || (rc = balloc(b, &b->d_cell_off, n + 1)) ||
      (rc = balloc(b, &b->d_flags, n)))
    return bail(rc);
  if (hipHostMalloc((void**)&b->h_flags, sizeof(int) * n, hipHostMallocDefault) != hipSuccess)
    return bail(fail(PLUM_E_NOMEM, "hipHostMalloc failed"));
  std::vector<int> off(n + 1, 0);
  for (int r = 0; r < n; r++) off[r + 1] = off[r] + replicas[r]->g.ncell;
  b->total_cells = off[n];
  if (hipMemcpy(b->d_cell_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "upload of the cell offsets failed"));
  for (plum_ctx* c : b->reps)
    if (hipStreamSynchronize(c->st) != hipSuccess) return bail(fail(PLUM_E_HIP, "replica stream failed"));
  if ((rc = batch_upload_state(b)) != PLUM_OK) return bail(rc);
  *out = b;
  return PLUM_OK;
}

int plum_batch_destroy(plum_batch* b) {
  if (!b) return PLUM_OK;
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_drop_graph(b);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
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
  if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return fail(PLUM_E_ARG, "blocks_per_cu must be 1..64");
  if (t->split_div < 16) return fail(PLUM_E_ARG, "split_div must be >= 16");
  if (t->sweeps_per_graph < 1 || t->sweeps_per_graph > 4096)
    return fail(PLUM_E_ARG, "sweeps_per_graph must be 1..4096");
  b->tune.blocks_per_cu = t->blocks_per_cu;
  b->tune.split_div = t->split_div;
  b->tune.sweeps_per_graph = t->sweeps_per_graph;
  CHK(batch_upload_state(b));
  return PLUM_OK;
}

int plum_batch_mc_run(plum_batch* b, const plum_mc_params* mcp, long long n_moves, int use_graph,
                      plum_batch_stats* total, plum_mc_stats* per) {
  if (!b || !mcp) return fail(PLUM_E_ARG, "null argument");
  if (n_moves < 0) return fail(PLUM_E_ARG, "negative move count");
  const int R = b->R;
  for (int r = 0; r < R; r++)
    if (b->reps[r]->pending) return fail(PLUM_E_STATE, "a replica has a pending trial move");
  std::vector<McDev> mc(R);
  for (int r = 0; r < R; r++) CHK(mc_setup(b->reps[r], &mcp[r], b->st, &mc[r]));
  HIPCHK(hipMemcpyAsync(b->dMC, mc.data(), sizeof(McDev) * R, hipMemcpyHostToDevice, b->st));
  CHK(batch_upload_state(b));
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
  }
  long long resync_every = 0;
  for (int r = 0; r < R; r++)
    if (mcp[r].resync_every > 0)
      resync_every = resync_every ? std::min<long long>(resync_every, mcp[r].resync_every) : mcp[r].resync_every;
  const auto t0 = std::chrono::steady_clock::now();
  long long done = 0, since_resync = 0;
  const int M = b->tune.sweeps_per_graph;
  while (done < n_moves) {
    const int m = (int)std::min<long long>(M, n_moves - done);
    CHK(batch_launch(b, m, use_graph != 0 && m == M));
    HIPCHK(hipMemcpyAsync(b->h_flags, b->d_flags, sizeof(int) * R, hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    done += m;
    since_resync += m;
    // a replica whose cell list overflowed skips its remaining sweeps; it is
    // finished alone once the batch is through
    if (resync_every > 0 && since_resync >= resync_every && done < n_moves) {
      CHK(batch_flush(b));
      HIPCHK(hipStreamSynchronize(b->st));
      for (plum_ctx* c : b->reps) {
        CHK(build_cells(c, c->g.cap));
        double f[8];
        CHK(run_full(c, 1, f));
      }
      CHK(batch_upload_state(b));
      since_resync = 0;
    }
  }
  CHK(batch_flush(b));
  HIPCHK(hipStreamSynchronize(b->st));
  const auto t1 = std::chrono::steady_clock::now();
  // replicas that overflowed: rebuild with more room, redo the skipped moves
  for (int r = 0; r < R; r++) {
    plum_ctx* c = b->reps[r];
    CHK(read_ctl(c, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const long long skipped = c->h_ctl->skipped - sk0[r];
    if (skipped <= 0 && !c->h_ctl->overflow) continue;
    int zero = 0;
    HIPCHK(hipMemcpyAsync(&c->d.ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
    CHK(build_cells(c, 2 * c->g.cap));
    if (skipped > 0) CHK(plum_mc_run(c, &mcp[r], skipped, nullptr));
  }
  CHK(batch_upload_state(b));
  plum_batch_stats st{};
  for (int r = 0; r < R; r++) {
    plum_ctx* c = b->reps[r];
    CHK(read_ctl(c, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const Ctl* h = c->h_ctl;
    const long long moves = ((long long)h->move_counter - (long long)m0[r]) - (h->skipped - sk0[r]);
    st.moves += moves;
    for (int k = 0; k < 5; k++) {
      st.attempted[k] += h->attempted[k] - att0[5 * (size_t)r + k];
      st.accepted[k] += h->accepted[k] - acc0[5 * (size_t)r + k];
    }
    if (h->skipped - sk0[r] > 0) st.overflow_replicas += 1;
    if (per) {
      plum_mc_stats& p = per[r];
      std::memset(&p, 0, sizeof(p));
      for (int k = 0; k < 5; k++) {
        p.attempted[k] = h->attempted[k] - att0[5 * (size_t)r + k];
        p.accepted[k] = h->accepted[k] - acc0[5 * (size_t)r + k];
      }
      p.moves = moves;
      to_terms(h->running, &p.totals);
    }
  }
  st.sweeps = n_moves;
  st.seconds = std::chrono::duration<double>(t1 - t0).count();
  if (per)
    for (int r = 0; r < R; r++) per[r].seconds = st.seconds;
  if (total) *total = st;
  return PLUM_OK;
}

// Times `moves_per_trial` sweeps for each candidate launch shape (work-groups
// per CU of k_compute_batch x candidate slots per real-space query block) and
// keeps the fastest.  The trial sweeps are ordinary Monte Carlo moves of the
// replicas.
int plum_batch_tune(plum_batch* b, const plum_mc_params* mcp, int moves_per_trial, plum_batch_tuning* best) {
  if (!b || !mcp) return fail(PLUM_E_ARG, "null argument");
  if (moves_per_trial < 1) return fail(PLUM_E_ARG, "moves_per_trial must be >= 1");
  const int bpc[] = {2, 4, 8, 16};
  const int div[] = {192, 384, 768, 1536};
  plum_batch_tuning keep = b->tune, win = b->tune;
  double best_s = 1e300;
  for (int i : bpc)
    for (int j : div) {
      plum_batch_tuning t = keep;
      t.blocks_per_cu = i;
      t.split_div = j;
      t.sweeps_per_graph = moves_per_trial;
      CHK(plum_batch_set_tuning(b, &t));
      plum_batch_stats st{};
      CHK(plum_batch_mc_run(b, mcp, moves_per_trial, 1, &st, nullptr));
      if (st.seconds < best_s) {
        best_s = st.seconds;
        win = t;
        win.sweep_us = 1e6 * st.seconds / moves_per_trial;
      }
    }
  win.sweeps_per_graph = keep.sweeps_per_graph;
  CHK(plum_batch_set_tuning(b, &win));
  b->tune.sweep_us = win.sweep_us;
  if (best) *best = b->tune;
  return PLUM_OK;
}

}  // extern "C"
