// Host side of plum_vp_sample (include/plum_gpu.h): one volume perturbation
// of the pressure estimator.  Included once, at the end of engine.hip.

struct VpHost {
  double* zs = nullptr;       // stretched z per bead
  double* ppart = nullptr;    // pair partials (Ewald, LJ) per block
  int4* kl_u = nullptr;       // half-space l vectors in the current or the stretched k list
  double *c_old = nullptr, *c_new = nullptr, *dc = nullptr;   // c_k of each box (0: not in its list), c' - c
  int nk_u = 0;
  double2* skpart = nullptr;  // S(k) and dS(k) partials [chunk][6][k]
  double *ekpart = nullptr, *out = nullptr;
  double* h_out = nullptr;    // pinned
  int npb = 0, nchunk = 0, nkb = 0;
};

void vp_free(VpHost* h) {
  if (!h) return;
  if (h->h_out) (void)hipHostFree(h->h_out);
  delete h;   // device buffers belong to the context's allocation list
}

namespace {

// The union of the current k list and the list of the box stretched along z
// by kDz: the same l grid (potential_ewald_coul.cc:86-113 keep repl_cell), z
// wavenumbers (l 2 kPi) / (Lz + kDz), the volume Lx Ly (Lz + kDz).  Where l
// is in both, c' - c = c expm1(ln(c'/c)), ln(c'/c) = -log1p(dz/Lz)
// - log1p((k'^2 - k^2)/k^2) - (k'^2 - k^2)/(4 alpha), k'^2 - k^2 =
// (kz' - kz)(kz' + kz), evaluated without the cancellation of c' - c.
int vp_setup(plum_ctx* c) {
  if (c->vp) return PLUM_OK;
  VpHost* h = new VpHost();
  c->vp = h;
  const Geo& g = c->g;
  const plum_info& info = c->info;
  std::vector<int4> kl;
  std::vector<double> co, cn, dcv;
  if (c->p.use_ewald) {
    const int* n3 = info.repl_cell;
    const double L = g.L[2], Ls = g.L[2] + kDz;
    const double pref_o = c->p.lB / (kPi * (g.L[0] * g.L[1] * L)) * (4 * kPi * kPi);
    const double pref_n = c->p.lB / (kPi * (g.L[0] * g.L[1] * Ls)) * (4 * kPi * kPi);
    const double a4 = 4 * c->p.alpha;
    for (int lx = -n3[0]; lx <= n3[0]; lx++)
      for (int ly = -n3[1]; ly <= n3[1]; ly++)
        for (int lz = -n3[2]; lz <= n3[2]; lz++) {
          const bool half = lx > 0 || (lx == 0 && (ly > 0 || (ly == 0 && lz > 0)));
          if (!half) continue;
          const double kx = (double)(lx * 2) * kPi / g.L[0];
          const double ky = (double)(ly * 2) * kPi / g.L[1];
          const double kz = (double)(lz * 2) * kPi / L;
          const double kzs = (double)(lz * 2) * kPi / Ls;
          const double k2 = kx * kx + ky * ky + kz * kz;
          const double k2s = kx * kx + ky * ky + kzs * kzs;
          const bool in_o = k2 > 0 && k2 <= info.repl_cutoff;
          const bool in_n = k2s > 0 && k2s <= info.repl_cutoff;
          if (!in_o && !in_n) continue;
          const double c_o = in_o ? pref_o * (exp(-k2 / a4) / k2) : 0.0;
          const double c_n = in_n ? pref_n * (exp(-k2s / a4) / k2s) : 0.0;
          double d = c_n - c_o;
          if (in_o && in_n) {
            // kz' - kz = (lz 2 kPi) (1/Ls - 1/L) = -(lz 2 kPi) kDz / (L Ls)
            const double dkz = -((double)(lz * 2) * kPi) * kDz / (L * Ls);
            const double dk2 = dkz * (kzs + kz);
            const double lr = -std::log1p(kDz / L) - std::log1p(dk2 / k2) - dk2 / a4;
            d = c_o * std::expm1(lr);
          }
          kl.push_back(make_int4(lx, ly, lz, 0));
          co.push_back(c_o);
          cn.push_back(c_n);
          dcv.push_back(d);
        }
  }
  h->nk_u = (int)kl.size();
  const long long n = g.n;
  const long long npair = n * (n + 1) / 2;
  h->npb = (int)std::max(1LL, std::min(8192LL, (npair + kBlock - 1) / kBlock));
  h->nchunk = (g.n + kVpChunk - 1) / kVpChunk;
  const int nk = std::max(h->nk_u, 1);
  h->nkb = (nk + kBlock - 1) / kBlock;
  CHK(dalloc(c, &h->zs, (size_t)g.n));
  CHK(dalloc(c, &h->ppart, 2 * (size_t)h->npb));
  CHK(dalloc(c, &h->kl_u, (size_t)nk));
  CHK(dalloc(c, &h->c_old, (size_t)nk));
  CHK(dalloc(c, &h->c_new, (size_t)nk));
  CHK(dalloc(c, &h->dc, (size_t)nk));
  CHK(dalloc(c, &h->skpart, (size_t)h->nchunk * 6 * nk));
  CHK(dalloc(c, &h->ekpart, (size_t)h->nkb));
  CHK(dalloc(c, &h->out, 6));
  if (hipHostMalloc((void**)&h->h_out, 6 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(PLUM_E_NOMEM, "hipHostMalloc failed");
  if (h->nk_u > 0) {
    HIPCHK(hipMemcpyAsync(h->kl_u, kl.data(), sizeof(int4) * kl.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(h->c_old, co.data(), sizeof(double) * co.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(h->c_new, cn.data(), sizeof(double) * cn.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(h->dc, dcv.data(), sizeof(double) * dcv.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  return PLUM_OK;
}

}  // namespace

extern "C" int plum_vp_sample(plum_ctx* c, int phantom, plum_vp* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  if (phantom < 0 || phantom > c->g.n_mols) return fail(PLUM_E_ARG, "phantom count out of range");
  for (int m = 0; m < phantom; m++)
    if (c->mol_start[m + 1] - c->mol_start[m] != 1)
      return fail(PLUM_E_ARG, "phantom surface sites must be single-bead molecules");
  if (c->self) HIPCHK(hipStreamSynchronize(c->self->st));
  if (!c->vp) {
    const int rc = vp_setup(c);
    if (rc != PLUM_OK) return rc;
  }
  VpHost* h = c->vp;
  const Geo& g = c->g;
  hipLaunchKernelGGL(k_vp_scale, dim3((g.n_mols + kBlock - 1) / kBlock), dim3(kBlock), 0, c->st, g, c->d, h->zs);
  const long long n = g.n;
  hipLaunchKernelGGL(k_vp_pairs, dim3(h->npb), dim3(kBlock), 0, c->st, g, c->d, h->zs, phantom, n * (n + 1) / 2,
                     h->ppart);
  int nkb = 0;
  if (g.use_ewald && h->nk_u > 0) {
    hipLaunchKernelGGL(k_vp_dsk, dim3((h->nk_u + kBlock - 1) / kBlock, h->nchunk), dim3(kBlock), 0, c->st, g, c->d,
                       h->kl_u, h->nk_u, h->zs, phantom, h->skpart);
    hipLaunchKernelGGL(k_vp_dsk_energy, dim3(h->nkb), dim3(kBlock), 0, c->st, h->skpart, h->nchunk, h->nk_u,
                       h->c_old, h->c_new, h->dc, h->ekpart);
    nkb = h->nkb;
  }
  hipLaunchKernelGGL(k_vp_small, dim3(1), dim3(kBlock), 0, c->st, g, c->d, h->zs, phantom, h->ppart, h->npb,
                     h->ekpart, nkb, h->out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h->h_out, h->out, 6 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  const double* o = h->h_out;
  out->el = o[0] + o[5];
  out->lj = o[1];
  out->wall = o[2];
  out->bond = o[3];
  out->dipole = o[4];
  out->dU = out->el + out->lj + out->wall + out->bond + out->dipole;
  out->n_mol = g.n_mols - phantom;
  out->vol = g.Lsim[0] * g.Lsim[1] * g.Lsim[2];
  out->area_dz = g.Lsim[0] * g.Lsim[1] * kDz;
  return PLUM_OK;
}
