// MI355X (gfx950) trial-move energy engine: host side of the C ABI declared
// in include/plum_gpu.h.  Kernels live in kernels_move.h / kernels_full.h.
//
// Data layout in HBM (one system per context):
//   beads (index order):  x,y,z current unwrapped FP64, q FP64, type, molecule
//   cell list:            per cell a high-water count and `cap` slots with
//                         slot-major copies of (wrapped x,y,z,q) as double4,
//                         a float4 copy relative to the cell origin for the
//                         FP32 path, the bead index (-1 = tombstone) and type
//   k space:              half-space k list (lx,ly,lz) int4, c_k FP64,
//                         running S(k) double2, pending dS(k) double2
//   pending move:         moved list (index, old/new position+charge) and the
//                         per-bead phase tables e^{i l theta_d}, |l| <= n_d
//   control block:        dipole moment, dipole lag state, running totals,
//                         the last dE per term, status flags, move counters
#include "kernels_batch.h"
#include "kernels_cbmc.h"
#include "kernels_grow.h"
#include "kernels_full.h"
#include "kernels_move.h"
#include "kernels_pressure.h"
#include "plum_gpu.h"
#include "surrogate.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

using namespace plum;

struct plum_batch;
struct SurHost;
struct CbmcHost;
struct VpHost;
void cbmc_free(CbmcHost* h);
void vp_free(VpHost* h);
extern "C" int plum_batch_destroy(plum_batch* batch);
void sur_free(SurHost* s);

struct plum_ctx {
  plum_params p{};
  Geo g{};
  plum_info info{};
  Dev d{};
  hipStream_t st = nullptr;
  int device = 0;        // HIP device the context lives on
  int maxm = 1;          // capacity of the moved list (largest molecule)
  int nkb = 0;           // blocks of the k-space delta kernel (64 k each)
  int nkb_full = 0;      // blocks of the full S(k) kernels (256 k each)
  int nmmb = 1;          // blocks of the moved-moved pair work
  int nchunk = 1;        // bead chunks of the S(k) build
  int chunk = 512;
  int nrealb = 1;        // blocks of the full real-space kernel
  int nsplit = 1;        // blocks per real-space query
  double cand = 0;       // expected candidate pair-images per query
  int cell_div = 2;      // cells per search radius and dimension
  int grid_gen = 0;      // bumped by every re-grid (batches re-upload their copies)
  int nkb_s = 0;         // k-space blocks of the speculative loop
  int nrun = 0;          // k-vector runs of the speculative k-space (Dev::krun)
  int xcap = 0;          // cross-term partial capacity
  plum_batch* self = nullptr;  // this system as a batch of one (device loop)
  SurHost* sur = nullptr;      // delayed-acceptance surrogate (null when off)
  CbmcHost* cbmc = nullptr;    // configurational-bias trial buffers
  VpHost* vp = nullptr;        // volume-perturbation pressure buffers
  int screen_cap = kMaxProp;   // proposal slots allocated
  double* epart_k = nullptr;
  std::vector<int> mol_start;
  std::vector<int> type_h;
  std::vector<double> q_h;
  char* h_in = nullptr;  // pinned staging of the move record
  Ctl* h_ctl = nullptr;  // pinned copy of the control block
  size_t in_bytes = 0, off_xyz = 0;
  char* d_in = nullptr;
  int drec_cap = 0;      // decision-record capacity allocated (plum_mc_record)
  hipEvent_t ev_idle = nullptr;  // marks the end of the work queued on st
  bool pending = false;
  bool profiling = false;
  std::vector<void*> dev_allocs;
  // kernel timing (profiling mode of plum_mc_run)
  std::vector<std::string> kname;
  std::vector<double> ksum_us;
  std::vector<long long> kcount;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                               \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess)                                                       \
      return fail(PLUM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)                      \
  do {                              \
    int rc_ = (x);                  \
    if (rc_ != PLUM_OK) return rc_; \
  } while (0)

template <typename T>
int dalloc(plum_ctx* c, T** ptr, size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * count);
  if (e != hipSuccess) return fail(PLUM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->dev_allocs.push_back(p);
  *ptr = static_cast<T*>(p);
  return PLUM_OK;
}

// potential_ewald_coul.cc:43-75: slab padding and the derived cutoffs.
void derive_cutoffs(const plum_params& p, plum_info& info) {
  double box[3] = {p.box[0], p.box[1], p.box[2]};
  if (p.dipole_correction) {
    const double min_padding = 150;
    if (box[2] * kDiCorrection > min_padding) box[2] = box[2] * kDiCorrection;
    else box[2] += min_padding;
  }
  for (int i = 0; i < 3; i++) info.box_e[i] = box[i];
  info.vol_e = box[0] * box[1] * box[2];
  if (!p.use_ewald) return;
  const double lB = p.lB, alpha = p.alpha;
  double real_cutoff = 1;
  while (0.5 * lB * 1 * 1 * erfc(sqrt(alpha) * real_cutoff) / real_cutoff > kEwaldCutoff)
    real_cutoff += 1;
  info.real_cutoff = real_cutoff;
  for (int i = 0; i < 3; i++) info.real_cell[i] = (int)ceil(real_cutoff / box[i]);
  double repl_cutoff = alpha;
  const double min_box_vol = pow(box[0], 3);
  while (lB * 1 * 1 / (2 * kPi * min_box_vol) * (4 * kPi * kPi) * exp(-repl_cutoff / (4 * alpha)) /
             repl_cutoff >
         kEwaldCutoff)
    repl_cutoff += alpha;
  info.repl_cutoff = repl_cutoff;
  for (int i = 0; i < 3; i++) info.repl_cell[i] = (int)ceil(sqrt(repl_cutoff) * box[i] / (2 * kPi));
}

// Same expressions as wrap_coord / cell_coord on the device.
double host_wrap(double x, double L, double invL) {
  double w = x - L * floor(x * invL);
  if (w >= L) w -= L;
  if (w < 0) w += L;
  return w;
}

int host_cell(double xw, double inv_h, int n) {
  int c = (int)floor(xw * inv_h);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// Builds the cell list from the current device positions (host counting
// sort in bead order: deterministic slot order) and uploads it.
int build_cells(plum_ctx* c, int min_cap) {
  const Geo& g = c->g;
  const int n = g.n;
  std::vector<double> x(n), y(n), z(n);
  HIPCHK(hipMemcpyAsync(x.data(), c->d.x, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(y.data(), c->d.y, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(z.data(), c->d.z, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  std::vector<int> cell(n), count(g.ncell, 0);
  std::vector<double> w(3 * (size_t)n);
  std::vector<int> cc3(3 * (size_t)n);
  for (int i = 0; i < n; i++) {
    const double p[3] = {x[i], y[i], z[i]};
    int cc = 0;
    for (int k = 0; k < 3; k++) {
      w[3 * (size_t)i + k] = host_wrap(p[k], g.L[k], g.inv_L[k]);
      cc3[3 * (size_t)i + k] = host_cell(w[3 * (size_t)i + k], g.inv_h[k], g.nc[k]);
      cc = cc * g.nc[k] + cc3[3 * (size_t)i + k];
    }
    cell[i] = cc;
    count[cc]++;
  }
  const int maxc = *std::max_element(count.begin(), count.end());
  const int cap = std::max(std::max(16, min_cap), ((maxc * 2 + 8 + 7) / 8) * 8);
  const size_t slots = (size_t)g.ncell * cap;
  if (cap != c->g.cap || c->d.cbead == nullptr) {
    CHK(dalloc(c, &c->d.cbead, slots));
    CHK(dalloc(c, &c->d.cpos, slots));
    CHK(dalloc(c, &c->d.cpos32, slots));
    CHK(dalloc(c, &c->d.ctype, slots));
    c->g.cap = cap;
    c->nrealb = (int)((slots + kBlock - 1) / kBlock);
    CHK(dalloc(c, &c->d.fpart, 2 * (size_t)c->nrealb));
  }
  std::vector<int> cbead(slots, -1), ctype(slots, 0), bslot(n);
  std::vector<double4> cpos(slots, make_double4(0, 0, 0, 0));
  std::vector<float4> cpos32(slots, make_float4(0, 0, 0, 0));
  std::fill(count.begin(), count.end(), 0);
  for (int i = 0; i < n; i++) {
    const int cc = cell[i];
    const int s = count[cc]++;
    const size_t si = (size_t)cc * cap + s;
    const double* wi = &w[3 * (size_t)i];
    const int* ci = &cc3[3 * (size_t)i];
    cbead[si] = i;
    ctype[si] = c->type_h[i];
    cpos[si] = make_double4(wi[0], wi[1], wi[2], c->q_h[i]);
    cpos32[si] = make_float4((float)(wi[0] - ci[0] * g.h[0]), (float)(wi[1] - ci[1] * g.h[1]),
                             (float)(wi[2] - ci[2] * g.h[2]), (float)c->q_h[i]);
    bslot[i] = s;
  }
  HIPCHK(hipMemcpyAsync(c->d.ccount, count.data(), sizeof(int) * g.ncell, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.cbead, cbead.data(), sizeof(int) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.ctype, ctype.data(), sizeof(int) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.cpos, cpos.data(), sizeof(double4) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.cpos32, cpos32.data(), sizeof(float4) * slots, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.bcell, cell.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d.bslot, bslot.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->info.cell_cap = cap;
  return PLUM_OK;
}

int split_for(double cand, int split_div) {
  return std::max(1, std::min(kMaxSplit, (int)std::lround(cand / split_div)));
}

// Cell grid of edge ~ rs / div in every dimension (at most 256 cells each).
void set_cell_grid(plum_ctx* c, int div) {
  Geo& g = c->g;
  for (int k = 0; k < 3; k++) {
    int nc = (int)floor(g.L[k] / (g.rs / div));
    nc = std::max(1, std::min(nc, 256));
    g.nc[k] = nc;
    g.h[k] = g.L[k] / nc;
    g.inv_h[k] = 1.0 / g.h[k];
    c->info.n_cells[k] = nc;
  }
  g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
  c->cell_div = div;
}

// Expected candidate slots of a real-space query (the cells within the search
// radius of a point) and the query split it gives.
void set_cand(plum_ctx* c) {
  const Geo& g = c->g;
  double cand = g.n / c->info.vol_e;
  for (int k = 0; k < 3; k++) cand *= 2 * g.rs + g.h[k];
  c->cand = cand;
  c->nsplit = split_for(cand, kDefaultSplitDiv);
}

// A new cell grid for the current positions: the cell list rebuilt from
// scratch (the old arrays stay allocated until the context is destroyed).
int regrid_cells(plum_ctx* c, int div) {
  if (div == c->cell_div) return PLUM_OK;
  HIPCHK(hipStreamSynchronize(c->st));
  const int old_ncell = c->g.ncell;
  set_cell_grid(c, div);
  if (c->g.ncell != old_ncell) CHK(dalloc(c, &c->d.ccount, c->g.ncell));
  c->d.cbead = nullptr;   // slot arrays sized for the new grid
  CHK(build_cells(c, 0));
  set_cand(c);
  c->grid_gen++;
  return PLUM_OK;
}

// From-scratch energy on the device; with set_running the running totals,
// dipole state and the per-molecule spring cache are reset to it.
int run_full(plum_ctx* c, int set_running, double out[8]) {
  const Geo& g = c->g;
  hipLaunchKernelGGL(k_full_real, dim3(c->nrealb), dim3(kBlock), 0, c->st, g, c->d);
  if (g.use_ewald && g.n_kh > 0) {
    hipLaunchKernelGGL(k_full_sk_part, dim3(c->nkb_full, c->nchunk), dim3(kBlock), 0, c->st, g, c->d,
                       c->chunk);
    hipLaunchKernelGGL(k_full_sk_reduce, dim3(c->nkb_full), dim3(kBlock), 0, c->st, g, c->d, c->nchunk,
                       c->epart_k);
  }
  hipLaunchKernelGGL(k_full_small, dim3(1), dim3(kBlock), 0, c->st, g, c->d, c->nrealb,
                     (g.use_ewald && g.n_kh > 0) ? c->nkb_full : 0, c->epart_k, set_running);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int k = 0; k < 8; k++) out[k] = c->h_ctl->full[k];
  return PLUM_OK;
}

void to_terms(const double t[8], plum_terms* o) {
  o->lj = t[T_LJ]; o->real = t[T_REAL]; o->kspace = t[T_KSPACE]; o->self = t[T_SELF];
  o->dipole = t[T_DIPOLE]; o->bond = t[T_BOND]; o->wall = t[T_WALL];
}

int validate(const plum_params* p, int n, const double* xyz, const double* q, const int* type,
             int n_mols, const int* ms) {
  if (!p || !xyz || !q || !type || !ms) return fail(PLUM_E_ARG, "null argument");
  if (n <= 0 || n_mols <= 0) return fail(PLUM_E_ARG, "empty system");
  if (p->npbc != 2 && p->npbc != 3)
    return fail(PLUM_E_ARG, "npbc must be 2 or 3 (simulation.cc:152-156)");
  if (p->npbc == 3 && p->dipole_correction)
    return fail(PLUM_E_UNSUPPORTED, "dipole correction with 3 periodic dimensions");
  if (!(p->beta > 0)) return fail(PLUM_E_ARG, "beta must be > 0 (simulation.cc:142-146)");
  for (int k = 0; k < 3; k++)
    if (!(p->box[k] > 0) || !std::isfinite(p->box[k])) return fail(PLUM_E_ARG, "box lengths must be > 0");
  if (p->n_types <= 0 || p->n_types > PLUM_MAX_TYPES) return fail(PLUM_E_ARG, "n_types out of range");
  if (p->use_ewald && !(p->alpha > 0 && p->lB >= 0)) return fail(PLUM_E_ARG, "Ewald alpha must be > 0");
  if (p->precision != PLUM_FP64 && p->precision != PLUM_FP32_PAIR)
    return fail(PLUM_E_ARG, "unknown precision");
  if (ms[0] != 0 || ms[n_mols] != n) return fail(PLUM_E_ARG, "mol_start must span 0..n_beads");
  for (int m = 0; m < n_mols; m++) {
    if (ms[m + 1] <= ms[m]) return fail(PLUM_E_ARG, "empty or unordered molecule");
    if (ms[m + 1] - ms[m] > kMaxChainDev)
      return fail(PLUM_E_UNSUPPORTED, "molecule longer than 2048 beads");
  }
  for (int i = 0; i < n; i++) {
    if (type[i] < 0 || type[i] >= p->n_types) return fail(PLUM_E_ARG, "bead type out of range");
    for (int k = 0; k < 3; k++)
      if (!std::isfinite(xyz[3 * (size_t)i + k])) return fail(PLUM_E_ARG, "non-finite coordinate");
    if (!std::isfinite(q[i])) return fail(PLUM_E_ARG, "non-finite charge");
  }
  return PLUM_OK;
}

int nmm_for(const plum_ctx*, int nm) { return nmm_blocks(nm); }

// Blocks per real-space query: about 1.5 candidate slots per thread.

// k-space and real-space work of one trial move in one launch, once the
// moved list (and, for the k space, the phase tables) exists on the device.
// nm is the exact count (host path) or the capacity (device loop; the
// kernels read the real count on the device).
void launch_compute(plum_ctx* c, int nm) {
  const Geo& g = c->g;
  const int nkb = (g.use_ewald && g.n_kh > 0) ? c->nkb : 0;
  const int nmm = nmm_for(c, nm);
  const int nreal = 2 * nm * c->nsplit;
  const int grid = nkb + nreal + nmm;
  if (c->p.precision == PLUM_FP32_PAIR)
    hipLaunchKernelGGL(k_compute<1>, dim3(grid), dim3(kBlock), 0, c->st, g, c->d, nkb, nreal, nmm, c->nsplit);
  else
    hipLaunchKernelGGL(k_compute<0>, dim3(grid), dim3(kBlock), 0, c->st, g, c->d, nkb, nreal, nmm, c->nsplit);
}

}  // namespace

extern "C" {

const char* plum_last_error(void) { return g_err.c_str(); }

int plum_create(const plum_params* params, int n, const double* xyz, const double* q, const int* type,
                int n_mols, const int* ms, plum_ctx** out) {
  if (!out) return fail(PLUM_E_ARG, "null output");
  *out = nullptr;
  CHK(validate(params, n, xyz, q, type, n_mols, ms));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PLUM_E_HIP, "no HIP device available");
  plum_ctx* c = new plum_ctx();
  c->p = *params;
  const plum_params& p = c->p;
  int rc = PLUM_OK;
  auto bail = [&](int code) {
    plum_destroy(c);
    return code;
  };
  plum_info& info = c->info;
  derive_cutoffs(p, info);
  Geo& g = c->g;
  for (int k = 0; k < 3; k++) {
    g.L[k] = info.box_e[k];
    g.Lsim[k] = p.box[k];
    g.inv_L[k] = 1.0 / g.L[k];
  }
  g.npbc = p.npbc;
  g.n = n;
  g.n_mols = n_mols;
  g.beta = p.beta;
  g.use_pair = p.use_pair;
  g.use_ewald = p.use_ewald;
  g.use_bond = p.use_bond;
  g.use_ext = p.use_ext;
  g.n_types = p.n_types;
  g.rc = p.use_ewald ? info.real_cutoff : 0;
  g.sqrt_alpha = p.use_ewald ? sqrt(p.alpha) : 0;
  g.lB = p.lB;
  g.self_coef = p.use_ewald ? -p.lB * sqrt(p.alpha / kPi) : 0;
  g.dip_coef = p.lB * 2 * kPi / (info.vol_e / 1.0);
  for (int k = 0; k < 3; k++) g.real_cell[k] = info.real_cell[k];
  g.k_bond = p.k_bond;
  g.r0 = p.r0;
  g.wall_cut = p.wall_cut;
  g.wall_r3_k213 = pow((1.0 / k213), 3);
  g.wall_r3_cut = pow((1.0 / p.wall_cut), 3);
  g.dipole = p.use_ewald && p.dipole_correction;
  g.dipole_exact = p.dipole_exact;

  // LJ pair table with the reference's expressions (potential_truncated_lj.cc:49-85)
  std::vector<LJPair> lj((size_t)p.n_types * p.n_types);
  std::vector<LJPair32> lj32(lj.size());
  double lj_max = 0;
  for (int a = 0; a < p.n_types; a++)
    for (int b = 0; b < p.n_types; b++) {
      LJPair& e = lj[(size_t)a * p.n_types + b];
      e.sigma = (p.lj_sigma[a] + p.lj_sigma[b]) / 2;
      e.eps = sqrt((p.lj_eps[a] * p.lj_eps[b]));
      if (p.lj_cutoff < 0) {
        const double r6_ref = pow((1.0 / k216), 6);
        e.shift = 4 * e.eps * (r6_ref * r6_ref - r6_ref);
        e.cut = k216 * e.sigma;
      } else {
        const double r6_ref = pow((e.sigma / p.lj_cutoff), 6);
        e.shift = 4 * e.eps * (r6_ref * r6_ref - r6_ref);
        e.cut = p.lj_cutoff;
      }
      LJPair32& f = lj32[(size_t)a * p.n_types + b];
      f.sigma2 = (float)(e.sigma * e.sigma);
      f.eps4 = (float)(4 * e.eps);
      f.cut = e.cut;
      f.shift = (float)e.shift;
      if (p.use_pair) lj_max = std::max(lj_max, e.cut);
    }
  for (int k = 0; k < p.npbc; k++)
    if (p.use_pair && !(p.box[k] > 2 * lj_max)) {
      delete c;
      return fail(PLUM_E_UNSUPPORTED, "periodic box shorter than twice the LJ cutoff");
    }
  std::vector<WallType> wall(p.n_types);
  for (int t = 0; t < p.n_types; t++) {
    wall[t].sigma = p.wall_sigma[t];
    wall[t].eps = p.wall_eps[t];
    wall[t].kind = p.type_kind[t];
    wall[t].pad = 0;
  }
  // FP32 path: pairs closer than 1.3 x the largest LJ sigma (the steep,
  // large close-contact terms), and pairs whose real-space term for unit
  // charges exceeds 3e-3 kT, are evaluated in FP64.  FP32 pair math carries
  // ~1e-7 of relative error per pair; summed over the ~10^5 pair images of a
  // long chain's move (old and new) the large near terms would leave ~1e-5 kT
  // of error in dE, the small far ones leave < 1e-6 kT.
  double sig_max = 0;
  for (int t = 0; t < p.n_types; t++) sig_max = std::max(sig_max, p.lj_sigma[t]);
  double r_u = 0;
  if (p.use_ewald && p.lB > 0) {
    const double sa = sqrt(p.alpha);
    r_u = 0.05;
    while (r_u < info.real_cutoff && p.lB * erfc(sa * r_u) / r_u > 3e-3) r_u += 0.05;
  }
  g.fp64_below = std::max(1.3 * std::max(sig_max, 0.0), r_u);
  // explicit pair sums need only the minimum image when no other image can
  // reach rc along a periodic edge (edges longer than 2 rc); a non-periodic z
  // is checked per pair (real_pair)
  g.single_image = p.use_ewald ? 1 : 0;
  for (int k = 0; k < p.npbc && p.use_ewald; k++)
    if (!(2 * g.rc * (1 + 1e-9) < g.L[k])) g.single_image = 0;
  g.rs = std::max(g.rc, lj_max);
  if (g.rs <= 0) g.rs = 1.0;
  g.rs2 = g.rs * g.rs;

  // cell grid: edge ~ rs / div
  set_cell_grid(c, p.cell_div > 0 ? p.cell_div : 2);

  // half-space k list (potential_ewald_coul.cc:78-115; k and -k paired)
  std::vector<int4> kl;
  std::vector<double> ck;
  info.n_k_full = 0;
  if (p.use_ewald) {
    const int* n3 = info.repl_cell;
    // (the z tables always hold l_z = 1, the step of the speculative
    // k-space's z recurrence)
    g.nkx = n3[0]; g.nky = n3[1]; g.nkz = std::max(n3[2], 1);
    const double pref0 = p.lB / (kPi * info.vol_e) * (4 * kPi * kPi);
    for (int lx = -n3[0]; lx <= n3[0]; lx++)
      for (int ly = -n3[1]; ly <= n3[1]; ly++)
        for (int lz = -n3[2]; lz <= n3[2]; lz++) {
          const double kx = (double)(lx * 2) * kPi / g.L[0];
          const double ky = (double)(ly * 2) * kPi / g.L[1];
          const double kz = (double)(lz * 2) * kPi / g.L[2];
          const double k2 = kx * kx + ky * ky + kz * kz;
          if (!(k2 > 0 && k2 <= info.repl_cutoff)) continue;
          info.n_k_full++;
          const bool half = lx > 0 || (lx == 0 && (ly > 0 || (ly == 0 && lz > 0)));
          if (!half) continue;
          kl.push_back(make_int4(lx, ly, lz, 0));
          ck.push_back(pref0 * (exp(-k2 / (4 * p.alpha)) / k2));
        }
    g.tab_w = (g.nkx + 1) + (g.nky + 1) + (g.nkz + 1);
  }
  g.n_kh = (int)kl.size();
  info.n_k_half = g.n_kh;
  // runs of up to kKRun k vectors with one (l_x, l_y) and consecutive l_z
  // (the list has l_z fastest), one lane each in the speculative k-space
  std::vector<int2> krun;
  for (int k = 0; k < g.n_kh;) {
    int len = 1;
    while (len < kKRun && k + len < g.n_kh && kl[k + len].x == kl[k].x && kl[k + len].y == kl[k].y &&
           kl[k + len].z == kl[k].z + len)
      len++;
    krun.push_back(make_int2(k, len));
    k += len;
  }
  c->nrun = (int)krun.size();
  c->nkb = std::max(1, (g.n_kh + kKPerBlock - 1) / kKPerBlock);
  c->nkb_full = std::max(1, (g.n_kh + kBlock - 1) / kBlock);

  c->mol_start.assign(ms, ms + n_mols + 1);
  c->type_h.assign(type, type + n);
  c->q_h.assign(q, q + n);
  int maxm = 1;
  for (int m = 0; m < n_mols; m++) maxm = std::max(maxm, ms[m + 1] - ms[m]);
  c->maxm = maxm;
  info.n_beads = n;
  info.n_mols = n_mols;
  info.max_mol_len = maxm;
  c->nmmb = nmm_blocks(maxm);
  c->nchunk = (n + c->chunk - 1) / c->chunk;

  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess) {
    delete c;
    return fail(PLUM_E_HIP, "hipStreamCreate failed");
  }
  if (hipEventCreateWithFlags(&c->ev_idle, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(c->st);
    delete c;
    return fail(PLUM_E_HIP, "hipEventCreate failed");
  }
  Dev& d = c->d;
  std::vector<double> hx(n), hy(n), hz(n);
  std::vector<int> mol_of(n);
  for (int i = 0; i < n; i++) {
    hx[i] = xyz[3 * (size_t)i]; hy[i] = xyz[3 * (size_t)i + 1]; hz[i] = xyz[3 * (size_t)i + 2];
  }
  for (int m = 0; m < n_mols; m++)
    for (int i = ms[m]; i < ms[m + 1]; i++) mol_of[i] = m;
#define ALLOC(ptr, cnt) if ((rc = dalloc(c, &ptr, cnt)) != PLUM_OK) return bail(rc)
  ALLOC(d.x, n); ALLOC(d.y, n); ALLOC(d.z, n); ALLOC(d.q, n);
  ALLOC(d.type, n); ALLOC(d.mol_of, n); ALLOC(d.mol_start, n_mols + 1);
  ALLOC(d.bcell, n); ALLOC(d.bslot, n); ALLOC(d.moved, n); ALLOC(d.mv_slot, n);
  ALLOC(d.ccount, g.ncell);
  // per-proposal slots (kMaxProp) for the speculative device loop; the host
  // path uses slot 0
  ALLOC(d.kl, g.n_kh); ALLOC(d.ck, g.n_kh); ALLOC(d.Sk, g.n_kh); ALLOC(d.dS, (size_t)kMaxProp * g.n_kh);
  ALLOC(d.tab, (size_t)kMaxProp * maxm * 2 * std::max(1, g.tab_w));
  ALLOC(d.lj, lj.size()); ALLOC(d.lj32, lj32.size()); ALLOC(d.wall, wall.size());
  ALLOC(d.bond_cur, n_mols);
  ALLOC(d.mv_old, (size_t)kMaxProp * maxm); ALLOC(d.mv_new, (size_t)kMaxProp * maxm);
  ALLOC(d.mv_type, (size_t)kMaxProp * maxm);
  ALLOC(d.recs, kMaxScreen); ALLOC(d.smv_idx, (size_t)kMaxProp * maxm);
  ALLOC(d.prop, kMaxScreen); ALLOC(d.plan, 1);
  ALLOC(d.krun, std::max(1, c->nrun));
  // erfc nodes {erfc(x_i), (2 / sqrt(pi)) exp(-x_i^2)}, x_i = i / kErfcPerUnit,
  // each rounded once from long double (erfc_nodes)
  std::vector<double2> erfc_h(kErfcNodes);
  for (int i = 0; i < kErfcNodes; i++) {
    const long double x = (long double)i / kErfcPerUnit;
    erfc_h[i] = make_double2((double)erfcl(x), (double)(2.0L / sqrtl(3.141592653589793238462643383279502884L) *
                                                        expl(-x * x)));
  }
  double2* erfc_d = nullptr;
  ALLOC(erfc_d, kErfcNodes);
  g.erfc_tab = erfc_d;
  std::vector<float2> erfc32_h(kErfcNodes);
  for (int i = 0; i < kErfcNodes; i++) erfc32_h[i] = make_float2((float)erfc_h[i].x, (float)erfc_h[i].y);
  float2* erfc32_d = nullptr;
  ALLOC(erfc32_d, kErfcNodes);
  g.erfc_tab32 = erfc32_d;
  c->nkb_s = (g.use_ewald && g.n_kh > 0) ? (c->nrun + kKSpec - 1) / kKSpec : 0;
  ALLOC(d.part_ks, (size_t)kKVals * std::max(1, c->nkb_s));
  c->xcap = kMaxPairs * cross_blocks(maxm, maxm);
  ALLOC(d.part_x, c->xcap);
  ALLOC(d.ctl, 1);
  set_cand(c);
  // room for any split (the batch autotuner may change it)
  ALLOC(d.part_k, c->nkb); ALLOC(d.part_r, (size_t)kMaxProp * 2 * maxm * kMaxQWave);
  ALLOC(d.part_mm, (size_t)kMaxProp * c->nmmb);
  ALLOC(d.sk_part, (size_t)c->nchunk * std::max(1, g.n_kh));
  ALLOC(c->epart_k, c->nkb_full);
  ALLOC(d.chain_ids, n_mols); ALLOC(d.ion_ids, n_mols);
  // move record: header + idx[maxm] + pad + xyz[3 maxm], contiguous
  c->off_xyz = sizeof(MoveRec) + sizeof(int) * maxm + ((sizeof(int) * maxm) % 8);
  c->in_bytes = c->off_xyz + sizeof(double) * 3 * maxm;
  {
    char* din = nullptr;
    ALLOC(din, c->in_bytes);
    c->d_in = din;
    d.rec = reinterpret_cast<MoveRec*>(din);
    d.mv_idx = reinterpret_cast<int*>(din + sizeof(MoveRec));
    d.mv_xyz = reinterpret_cast<double*>(din + c->off_xyz);
  }
#undef ALLOC
  if (hipHostMalloc((void**)&c->h_in, c->in_bytes, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess)
    return bail(fail(PLUM_E_NOMEM, "hipHostMalloc failed"));
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return PLUM_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->st));
    return PLUM_OK;
  };
  if ((rc = up(d.x, hx.data(), sizeof(double) * n)) || (rc = up(d.y, hy.data(), sizeof(double) * n)) ||
      (rc = up(d.z, hz.data(), sizeof(double) * n)) || (rc = up(d.q, q, sizeof(double) * n)) ||
      (rc = up(d.type, type, sizeof(int) * n)) || (rc = up(d.mol_of, mol_of.data(), sizeof(int) * n)) ||
      (rc = up(d.mol_start, ms, sizeof(int) * (n_mols + 1))) ||
      (rc = up(d.kl, kl.data(), sizeof(int4) * kl.size())) ||
      (rc = up(d.krun, krun.data(), sizeof(int2) * krun.size())) ||
      (rc = up(d.ck, ck.data(), sizeof(double) * ck.size())) ||
      (rc = up(d.lj, lj.data(), sizeof(LJPair) * lj.size())) ||
      (rc = up(d.lj32, lj32.data(), sizeof(LJPair32) * lj32.size())) ||
      (rc = up(d.wall, wall.data(), sizeof(WallType) * wall.size())) ||
      (rc = up(erfc_d, erfc_h.data(), sizeof(double2) * kErfcNodes)) ||
      (rc = up(erfc32_d, erfc32_h.data(), sizeof(float2) * kErfcNodes)))
    return bail(rc);
  if (hipMemsetAsync(d.moved, 0, n, c->st) != hipSuccess ||
      hipMemsetAsync(d.mv_slot, 0xff, sizeof(int) * n, c->st) != hipSuccess ||
      hipMemsetAsync(d.ctl, 0, sizeof(Ctl), c->st) != hipSuccess ||
      hipMemsetAsync(d.plan, 0, sizeof(SpecPlan), c->st) != hipSuccess ||
      hipMemsetAsync(d.recs, 0, sizeof(MoveRec) * kMaxScreen, c->st) != hipSuccess ||
      hipMemsetAsync(c->d_in, 0, c->in_bytes, c->st) != hipSuccess)
    return bail(fail(PLUM_E_HIP, "hipMemsetAsync failed"));
  if ((rc = build_cells(c, 0)) != PLUM_OK) return bail(rc);
  double f[8];
  if ((rc = run_full(c, 1, f)) != PLUM_OK) return bail(rc);
  *out = c;
  return PLUM_OK;
}

int plum_destroy(plum_ctx* c) {
  if (!c) return PLUM_OK;
  // Teardown order: drain the context stream first (a rollback's commit
  // kernel may still be queued on it), then the device-loop batch (it drains
  // its own stream before destroying its graph), then the growth graph and
  // the surrogate, and only then the buffers those graphs and kernels read.
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->self) plum_batch_destroy(c->self);
  c->self = nullptr;
  if (c->cbmc) cbmc_free(c->cbmc);
  c->cbmc = nullptr;
  if (c->sur) sur_free(c->sur);
  c->sur = nullptr;
  if (c->vp) vp_free(c->vp);
  c->vp = nullptr;
  for (void* p : c->dev_allocs) (void)hipFree(p);
  if (c->h_in) (void)hipHostFree(c->h_in);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  if (c->ev_idle) (void)hipEventDestroy(c->ev_idle);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return PLUM_OK;
}

int plum_get_info(const plum_ctx* c, plum_info* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  *out = c->info;
  return PLUM_OK;
}

int plum_energy_full(plum_ctx* c, plum_terms* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  double f[8];
  CHK(run_full(c, 0, f));
  to_terms(f, out);
  return PLUM_OK;
}

int plum_energy_difference(plum_ctx* c, int mol, int n_moved, const int* idx, const double* trial_xyz,
                           plum_terms* dE, double* total) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (c->pending) return fail(PLUM_E_STATE, "previous trial move not finalized");
  if (mol < 0 || mol >= c->g.n_mols) return fail(PLUM_E_ARG, "molecule index out of range");
  const int a = c->mol_start[mol], b = c->mol_start[mol + 1];
  if (n_moved < 0 || n_moved > b - a) return fail(PLUM_E_ARG, "n_moved out of range");
  if (n_moved > 0 && (!idx || !trial_xyz)) return fail(PLUM_E_ARG, "null moved list");
  std::vector<char> seen(b - a, 0);
  for (int k = 0; k < n_moved; k++) {
    if (idx[k] < a || idx[k] >= b) return fail(PLUM_E_ARG, "moved bead not in the molecule");
    if (seen[idx[k] - a]++) return fail(PLUM_E_ARG, "moved bead listed twice");
    for (int e = 0; e < 3; e++)
      if (!std::isfinite(trial_xyz[3 * k + e])) return fail(PLUM_E_ARG, "non-finite trial coordinate");
  }
  MoveRec* r = reinterpret_cast<MoveRec*>(c->h_in);
  r->n_moved = n_moved;
  r->mol = mol;
  r->move_type = -1;
  r->valid = 1;
  int lo = b, hi = a - 1;
  for (int k = 0; k < n_moved; k++) {
    lo = std::min(lo, idx[k]);
    hi = std::max(hi, idx[k]);
  }
  r->contig = (n_moved > 0 && hi - lo + 1 == n_moved) ? 1 : 0;
  r->lo = lo;
  r->hi = hi;
  std::memcpy(c->h_in + sizeof(MoveRec), idx, sizeof(int) * n_moved);
  std::memcpy(c->h_in + c->off_xyz, trial_xyz, sizeof(double) * 3 * n_moved);
  HIPCHK(hipMemcpyAsync(c->d_in, c->h_in, sizeof(MoveRec) + sizeof(int) * n_moved, hipMemcpyHostToDevice,
                        c->st));
  if (n_moved > 0)
    HIPCHK(hipMemcpyAsync(c->d_in + c->off_xyz, c->h_in + c->off_xyz, sizeof(double) * 3 * n_moved,
                          hipMemcpyHostToDevice, c->st));
  const Geo& g = c->g;
  hipLaunchKernelGGL(k_host_moves, dim3(std::max(1, (n_moved + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->st,
                     g, c->d);
  if (g.use_ewald && g.n_kh > 0 && n_moved > 0) {
    const int tw = n_moved * 2 * std::max(1, g.tab_w);
    hipLaunchKernelGGL(k_tables, dim3(std::max(1, std::min(2048, (tw + kBlock - 1) / kBlock))), dim3(kBlock), 0,
                       c->st, g, c->d);
  }
  launch_compute(c, n_moved);
  hipLaunchKernelGGL(k_epilogue, dim3(1), dim3(kBlock), 0, c->st, g, c->d,
                     (g.use_ewald && g.n_kh > 0) ? c->nkb : 0, 2 * n_moved * c->nsplit, nmm_for(c, n_moved));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->pending = true;
  if (dE) to_terms(c->h_ctl->terms, dE);
  if (total) *total = c->h_ctl->total;
  return PLUM_OK;
}

int plum_finalize(plum_ctx* c, int accept) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (!c->pending) return fail(PLUM_E_STATE, "no pending trial move");
  const Geo& g = c->g;
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(kBlock), 0, c->st, g, c->d, accept ? 1 : 0);
  if (g.use_ewald && g.n_kh > 0 && accept)
    hipLaunchKernelGGL(k_commit_sk, dim3(c->nkb_full), dim3(kBlock), 0, c->st, g, c->d);
  HIPCHK(hipGetLastError());
  c->pending = false;
  if (accept) {
    HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->h_ctl->overflow) {
      // a cell ran out of slots: rebuild with twice the capacity
      int zero = 0;
      HIPCHK(hipMemcpyAsync(&c->d.ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
      CHK(build_cells(c, 2 * c->g.cap));
    }
  }
  return PLUM_OK;
}

int plum_totals(plum_ctx* c, plum_terms* out) {
  if (!c || !out) return fail(PLUM_E_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  to_terms(c->h_ctl->running, out);
  return PLUM_OK;
}

int plum_resync(plum_ctx* c, plum_terms* drift) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  if (c->pending) return fail(PLUM_E_STATE, "a trial move is pending finalize");
  HIPCHK(hipMemcpyAsync(c->h_ctl, c->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  double before[8];
  for (int k = 0; k < 8; k++) before[k] = c->h_ctl->running[k];
  int zero = 0;
  HIPCHK(hipMemcpyAsync(&c->d.ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
  CHK(build_cells(c, c->g.cap));
  double f[8];
  CHK(run_full(c, 1, f));
  if (drift) {
    double dd[8];
    for (int k = 0; k < 8; k++) dd[k] = before[k] - f[k];
    to_terms(dd, drift);
  }
  return PLUM_OK;
}

int plum_get_positions(plum_ctx* c, double* xyz) {
  if (!c || !xyz) return fail(PLUM_E_ARG, "null argument");
  const int n = c->g.n;
  std::vector<double> x(n), y(n), z(n);
  HIPCHK(hipMemcpyAsync(x.data(), c->d.x, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(y.data(), c->d.y, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(z.data(), c->d.z, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int i = 0; i < n; i++) {
    xyz[3 * (size_t)i] = x[i]; xyz[3 * (size_t)i + 1] = y[i]; xyz[3 * (size_t)i + 2] = z[i];
  }
  return PLUM_OK;
}

int plum_set_profiling(plum_ctx* c, int on) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  c->profiling = on != 0;
  return PLUM_OK;
}

int plum_kernel_times(plum_ctx* c, int cap, const char** names, double* avg_us, long long* count) {
  if (!c) return fail(PLUM_E_ARG, "null context");
  int n = 0;
  for (size_t i = 0; i < c->kname.size() && n < cap; i++) {
    if (!c->kcount[i]) continue;   // kernels that did not run
    if (names) names[n] = c->kname[i].c_str();
    if (avg_us) avg_us[n] = c->ksum_us[i] / c->kcount[i];
    if (count) count[n] = c->kcount[i];
    n++;
  }
  return n;
}

}  // extern "C"

#include "batch_api.h"
#include "surrogate_api.h"
#include "cbmc_api.h"
#include "noalloc_api.h"
#include "pressure_api.h"
