// Host side of the delayed-acceptance surrogate (plum_surrogate_* in
// include/plum_gpu.h).  Included once, at the end of engine.hip, after
// batch_api.h (SurHost and the batch loop that launches the surrogate).

namespace {

template <typename T>
int salloc(SurHost* s, T** ptr, size_t count) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(1, count));
  if (e != hipSuccess) return fail(PLUM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->allocs.push_back(p);
  *ptr = static_cast<T*>(p);
  return PLUM_OK;
}

unsigned nblocks(size_t n) { return (unsigned)std::max<size_t>(1, (n + kBlock - 1) / kBlock); }

// Training workspace for `samples` samples.
int sur_workspace(plum_ctx* c, int samples) {
  SurHost* s = c->sur;
  if (samples <= s->tb_cap) return PLUM_OK;
  const size_t M = (size_t)samples * 2 * c->maxm;
  CHK(salloc(s, &s->X, M * kDescr));
  CHK(salloc(s, &s->H1, M * kHid));
  CHK(salloc(s, &s->H2, M * kHid));
  CHK(salloc(s, &s->H3, M * kHid));
  CHK(salloc(s, &s->D3, M * kHid));
  CHK(salloc(s, &s->D2, M * kHid));
  CHK(salloc(s, &s->D1, M * kHid));
  CHK(salloc(s, &s->F, M));
  CHK(salloc(s, &s->G, M));
  CHK(salloc(s, &s->T, samples));
  CHK(salloc(s, &s->N, samples));
  CHK(salloc(s, &s->part, ((M + kAtbRows - 1) / kAtbRows) * (size_t)(kHid * kHid + kHid)));
  s->tb_cap = samples;
  return PLUM_OK;
}

// Forward (FP32) of the M = samples * 2 maxm rows in s->X.
void sur_forward32(plum_ctx* c, hipStream_t st, size_t M) {
  SurHost* s = c->sur;
  const float* w = s->w32;
  const int m = (int)M;
  hipLaunchKernelGGL(k_gemm_ab, dim3(nblocks(M * kHid)), dim3(kBlock), 0, st, s->X, w + kOffW1, w + kOffB1, s->H1, m,
                     kHid, kDescr, 1);
  hipLaunchKernelGGL(k_gemm_ab, dim3(nblocks(M * kHid)), dim3(kBlock), 0, st, s->H1, w + kOffW2, w + kOffB2, s->H2,
                     m, kHid, kHid, 1);
  hipLaunchKernelGGL(k_gemm_ab, dim3(nblocks(M * kHid)), dim3(kBlock), 0, st, s->H2, w + kOffW3, w + kOffB3, s->H3,
                     m, kHid, kHid, 1);
  hipLaunchKernelGGL(k_gemm_ab, dim3(nblocks(M)), dim3(kBlock), 0, st, s->H3, w + kOffW4, w + kOffB4, s->F, m, 1,
                     kHid, 0);
}

// dW = A^T B (and the bias gradient, column sums of B) over M rows.
void sur_atb(SurHost* s, hipStream_t st, const float* A, const float* B, float* Gm, float* bias_g, int M, int K1,
             int K2) {
  const int chunks = (M + kAtbRows - 1) / kAtbRows;
  const int nout = K1 * K2 + K2;
  hipLaunchKernelGGL(k_gemm_atb_part, dim3(nblocks(nout), chunks), dim3(kBlock), 0, st, A, B, s->part, M, K1, K2);
  hipLaunchKernelGGL(k_gemm_atb_sum, dim3(nblocks(nout)), dim3(kBlock), 0, st, s->part, chunks, K1, K2, Gm, bias_g);
}

// Loss and gradient of the samples in s->X / s->T / s->N, then (update) one
// Adam step and the BF16 repack.
int sur_train_step(plum_ctx* c, hipStream_t st, int samples, bool gather, bool update, bool sync) {
  SurHost* s = c->sur;
  CHK(sur_workspace(c, samples));
  const size_t M = (size_t)samples * 2 * c->maxm;
  if (gather)
    hipLaunchKernelGGL(k_gather_samples, dim3(16, samples), dim3(kBlock), 0, st, s->tr_x, s->tr_target, s->tr_n,
                       s->tr_ctl, samples, c->maxm, s->X, s->T, s->N);
  sur_forward32(c, st, M);
  hipLaunchKernelGGL(k_train_loss, dim3(1), dim3(kBlock), 0, st, s->F, s->T, s->N, samples, c->maxm,
                     s->p.target_clip, s->G, s->loss);
  float* g = s->grad;
  const float* w = s->w32;
  const int m = (int)M;
  sur_atb(s, st, s->H3, s->G, g + kOffW4, g + kOffB4, m, kHid, 1);
  hipLaunchKernelGGL(k_out_delta, dim3(nblocks(M * kHid)), dim3(kBlock), 0, st, s->G, w + kOffW4, s->H3, s->D3, m);
  sur_atb(s, st, s->H2, s->D3, g + kOffW3, g + kOffB3, m, kHid, kHid);
  hipLaunchKernelGGL(k_gemm_abt_dtanh, dim3(nblocks(M * kHid)), dim3(kBlock), 0, st, s->D3, w + kOffW3, s->H2, s->D2,
                     m, kHid, kHid);
  sur_atb(s, st, s->H1, s->D2, g + kOffW2, g + kOffB2, m, kHid, kHid);
  hipLaunchKernelGGL(k_gemm_abt_dtanh, dim3(nblocks(M * kHid)), dim3(kBlock), 0, st, s->D2, w + kOffW2, s->H1, s->D1,
                     m, kHid, kHid);
  sur_atb(s, st, s->X, s->D1, g + kOffW1, g + kOffB1, m, kDescr, kHid);
  if (update) {
    s->step++;
    hipLaunchKernelGGL(k_adam, dim3(nblocks(kNParams)), dim3(kBlock), 0, st, s->w32, s->grad, s->m1, s->m2, kNParams,
                       (float)s->p.lr, s->step);
    hipLaunchKernelGGL(k_pack_bf16, dim3(nblocks(kHid * kHid)), dim3(kBlock), 0, st, s->w32, s->net);
    s->train_steps++;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->h_loss, s->loss, sizeof(double), hipMemcpyDeviceToHost, st));
  if (sync) {
    HIPCHK(hipStreamSynchronize(st));
    s->last_loss = *s->h_loss;
  }
  return PLUM_OK;
}

hipStream_t sur_stream(plum_ctx* c) { return c->self ? c->self->st : c->st; }

}  // namespace

extern "C" {

int plum_max_mol_len(const plum_ctx* c) { return c ? c->maxm : PLUM_E_ARG; }

int plum_surrogate_enable(plum_ctx* c, const plum_surrogate_params* p) {
  if (!c || !p) return fail(PLUM_E_ARG, "null argument");
  if (p->screen < 1 || p->screen > kMaxScreen) return fail(PLUM_E_ARG, "screen must be 1..32");
  if (p->fill < 0 || p->fill > kMaxProp) return fail(PLUM_E_ARG, "fill must be 0..16 (the exact slots)");
  if (p->train_every < 0 || p->train_samples < 1 || p->ring < p->train_samples)
    return fail(PLUM_E_ARG, "need train_samples >= 1 and ring >= train_samples");
  if (!(p->r_d > kRbfMin + 0.5) || !(p->lr > 0) || !(p->target_clip > 0))
    return fail(PLUM_E_ARG, "need r_d > 1.2, lr > 0, target_clip > 0");
  if (p->r_d > c->g.rs) return fail(PLUM_E_ARG, "descriptor radius beyond the cell-list search radius");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  if (c->sur) {
    sur_free(c->sur);
    c->sur = nullptr;
  }
  // proposal slots for the screened batch
  CHK(ensure_screen_slots(c));
  SurHost* s = new SurHost();
  s->p = *p;
  c->sur = s;
  auto bail = [&](int code) {
    sur_free(c->sur);
    c->sur = nullptr;
    return code;
  };
  int rc = PLUM_OK;
  const size_t rows = (size_t)kMaxScreen * 2 * c->maxm;
  if ((rc = salloc(s, &s->net, 1)) || (rc = salloc(s, &s->w32, kNParams)) || (rc = salloc(s, &s->m1, kNParams)) ||
      (rc = salloc(s, &s->m2, kNParams)) || (rc = salloc(s, &s->grad, kNParams)) ||
      (rc = salloc(s, &s->sur_x, rows * kDescr)) || (rc = salloc(s, &s->sur_f, rows)) ||
      (rc = salloc(s, &s->tr_x, (size_t)p->ring * 2 * c->maxm * kDescr)) ||
      (rc = salloc(s, &s->tr_target, p->ring)) || (rc = salloc(s, &s->tr_n, p->ring)) ||
      (rc = salloc(s, &s->tr_ctl, 2)) || (rc = salloc(s, &s->loss, 1)))
    return bail(rc);
  if (hipHostMalloc((void**)&s->h_ctl, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->h_loss, sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(fail(PLUM_E_NOMEM, "hipHostMalloc failed"));
  *s->h_loss = 0;
  // weights: uniform(-a, a), a = sqrt(6 / (fan_in + fan_out)) (Glorot &
  // Bengio 2010); output row scaled by 0.1; biases zero
  std::vector<float> w(kNParams, 0.f);
  std::mt19937_64 gen(p->seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  auto fill = [&](int off, int fin, int fout, double scale) {
    const double a = scale * std::sqrt(6.0 / (fin + fout));
    for (int i = 0; i < fin * fout; i++) w[off + i] = (float)(a * u(gen));
  };
  fill(kOffW1, kDescr, kHid, 1.0);
  fill(kOffW2, kHid, kHid, 1.0);
  fill(kOffW3, kHid, kHid, 1.0);
  fill(kOffW4, kHid, 1, 0.1);
  const hipStream_t st = sur_stream(c);
  const int ctl[2] = {0, p->ring};
  if (hipMemcpyAsync(s->w32, w.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(s->m1, 0, sizeof(float) * kNParams, st) != hipSuccess ||
      hipMemsetAsync(s->m2, 0, sizeof(float) * kNParams, st) != hipSuccess ||
      hipMemsetAsync(s->sur_x, 0, sizeof(float) * rows * kDescr, st) != hipSuccess ||
      hipMemsetAsync(s->sur_f, 0, sizeof(float) * rows, st) != hipSuccess ||
      hipMemcpyAsync(s->tr_ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice, st) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "surrogate upload failed"));
  hipLaunchKernelGGL(k_pack_bf16, dim3(nblocks(kHid * kHid)), dim3(kBlock), 0, st, s->w32, s->net);
  if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(PLUM_E_HIP, "surrogate initialisation failed"));
  c->d.sur_x = s->sur_x;
  c->d.sur_f = s->sur_f;
  c->d.tr_x = s->tr_x;
  c->d.tr_target = s->tr_target;
  c->d.tr_n = s->tr_n;
  c->d.tr_ctl = s->tr_ctl;
  s->on = true;
  if (!c->self) {
    plum_ctx* one[1] = {c};
    if ((rc = batch_new(one, 1, &c->self)) != PLUM_OK) return bail(rc);
  }
  c->self->tune.pmax = p->screen;
  c->self->tune.batches_per_graph = p->train_every > 0 ? p->train_every : 32;
  batch_drop_graph(c->self);
  if ((rc = batch_upload_state(c->self)) != PLUM_OK) return bail(rc);
  return PLUM_OK;
}

int plum_surrogate_disable(plum_ctx* c) {
  if (!c) return fail(PLUM_E_ARG, "null argument");
  if (c->sur) {
    sur_free(c->sur);
    c->sur = nullptr;
  }
  c->d.sur_x = c->d.sur_f = c->d.tr_x = nullptr;
  c->d.tr_target = nullptr;
  c->d.tr_n = c->d.tr_ctl = nullptr;
  if (c->self) {
    c->self->tune.pmax = std::min(c->self->tune.pmax, kMaxProp);
    c->self->tune.batches_per_graph = 32;
    batch_drop_graph(c->self);
    CHK(batch_upload_state(c->self));
  }
  return PLUM_OK;
}

int plum_surrogate_stats_get(plum_ctx* c, plum_surrogate_stats* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  const hipStream_t st = sur_stream(c);
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(c->sur->h_ctl, c->sur->tr_ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  c->sur->last_loss = *c->sur->h_loss;
  out->s1_rejected = c->h_ctl->s1_rejected;
  out->s1_passed = c->h_ctl->s1_passed;
  out->train_evals = c->h_ctl->train_evals;
  out->stale_cuts = c->h_ctl->stale_cuts;
  out->samples = c->sur->h_ctl[0];
  out->train_steps = c->sur->train_steps;
  out->last_loss = c->sur->last_loss;
  return PLUM_OK;
}

int plum_surrogate_get_weights(plum_ctx* c, float* w, int n) {
  if (!c || !w) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  if (n != kNParams) return fail(PLUM_E_ARG, "weight count must be PLUM_SURROGATE_PARAMS");
  const hipStream_t st = sur_stream(c);
  HIPCHK(hipMemcpyAsync(w, c->sur->w32, sizeof(float) * kNParams, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return PLUM_OK;
}

int plum_surrogate_set_weights(plum_ctx* c, const float* w, int n) {
  if (!c || !w) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  if (n != kNParams) return fail(PLUM_E_ARG, "weight count must be PLUM_SURROGATE_PARAMS");
  for (int i = 0; i < n; i++)
    if (!std::isfinite(w[i])) return fail(PLUM_E_ARG, "non-finite weight");
  const hipStream_t st = sur_stream(c);
  HIPCHK(hipMemcpyAsync(c->sur->w32, w, sizeof(float) * kNParams, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pack_bf16, dim3(nblocks(kHid * kHid)), dim3(kBlock), 0, st, c->sur->w32, c->sur->net);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return PLUM_OK;
}

int plum_surrogate_train(plum_ctx* c, int steps, double* loss) {
  if (!c) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  if (steps < 0) return fail(PLUM_E_ARG, "negative step count");
  const hipStream_t st = sur_stream(c);
  HIPCHK(hipMemcpyAsync(c->sur->h_ctl, c->sur->tr_ctl, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int have = std::min(c->sur->h_ctl[0], c->sur->p.ring);
  const int samples = std::min(have, c->sur->p.train_samples);
  if (samples < 1) return fail(PLUM_E_STATE, "no training samples collected yet");
  for (int i = 0; i < steps; i++) CHK(sur_train_step(c, st, samples, true, true, i == steps - 1));
  if (loss) *loss = c->sur->last_loss;
  return PLUM_OK;
}

int plum_surrogate_forward(plum_ctx* c, const float* x, int rows, int bf16, float* f) {
  if (!c || !x || !f) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  if (rows < 1) return fail(PLUM_E_ARG, "rows must be >= 1");
  SurHost* s = c->sur;
  const hipStream_t st = sur_stream(c);
  const int per = 2 * c->maxm;
  CHK(sur_workspace(c, (rows + per - 1) / per));
  HIPCHK(hipMemcpyAsync(s->X, x, sizeof(float) * rows * kDescr, hipMemcpyHostToDevice, st));
  if (bf16)
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::max(1, std::min(256, (rows + 127) / 128))), dim3(kBlock), 0, st, s->net,
                       s->X, rows, s->F, rows, rows);
  else
    sur_forward32(c, st, rows);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(f, s->F, sizeof(float) * rows, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return PLUM_OK;
}

int plum_surrogate_gradient(plum_ctx* c, const float* x, const double* target, const int* n, int samples,
                            double* loss, float* grad) {
  if (!c || !x || !target || !n || !grad) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  if (samples < 1) return fail(PLUM_E_ARG, "samples must be >= 1");
  for (int i = 0; i < samples; i++)
    if (n[i] < 0 || n[i] > c->maxm) return fail(PLUM_E_ARG, "bead count out of range");
  SurHost* s = c->sur;
  const hipStream_t st = sur_stream(c);
  CHK(sur_workspace(c, samples));
  const size_t M = (size_t)samples * 2 * c->maxm;
  HIPCHK(hipMemcpyAsync(s->X, x, sizeof(float) * M * kDescr, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->T, target, sizeof(double) * samples, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->N, n, sizeof(int) * samples, hipMemcpyHostToDevice, st));
  CHK(sur_train_step(c, st, samples, false, false, true));
  HIPCHK(hipMemcpyAsync(grad, s->grad, sizeof(float) * kNParams, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (loss) *loss = s->last_loss;
  return PLUM_OK;
}

int plum_surrogate_debug_slot(plum_ctx* c, int slot, float* rows, double* pos, int* n) {
  if (!c || !rows || !pos || !n) return fail(PLUM_E_ARG, "null argument");
  if (!c->sur) return fail(PLUM_E_STATE, "surrogate not enabled");
  if (slot < 0 || slot >= kMaxScreen) return fail(PLUM_E_ARG, "slot out of range");
  const hipStream_t st = sur_stream(c);
  const int M = c->maxm;
  MoveRec rec;
  HIPCHK(hipMemcpyAsync(&rec, c->d.recs + slot, sizeof(MoveRec), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(rows, c->sur->sur_x + (size_t)2 * slot * M * kDescr, sizeof(float) * 2 * M * kDescr,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(pos, c->d.mv_old + (size_t)slot * M, sizeof(double4) * M, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(pos + 4 * (size_t)M, c->d.mv_new + (size_t)slot * M, sizeof(double4) * M,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *n = rec.n_moved;
  return PLUM_OK;
}

}  // extern "C"
