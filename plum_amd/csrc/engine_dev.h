// Device-side building blocks of the trial-move energy engine (gfx950).
// Shared by the kernels in engine.hip; no host code here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plum {

// utilities/constants.h:11-62 of the reference.
constexpr double kVeryLargeEnergy = 1E+8;
constexpr double kPi = 3.14159265359;
constexpr double kEwaldCutoff = 1E-7;
constexpr double k213 = 1.25992104989;
constexpr double k216 = 1.12246204831;
constexpr int kDiCorrection = 3;

constexpr int kBlock = 256;      // work-group size of the wide kernels
constexpr int kXcds = 8;         // XCDs of an MI355X; workgroup i is dispatched to XCD i % 8
constexpr int kWave = 64;        // CDNA wavefront
constexpr int kMaxQueryCells = 1024;

enum Term { T_LJ = 0, T_REAL, T_KSPACE, T_SELF, T_DIPOLE, T_BOND, T_WALL, NTERMS };

// Geometry and potential scalars, passed by value to every kernel.
struct Geo {
  double L[3];       // Ewald box: period of the real-space image sum in x,y,z
  double Lsim[3];    // simulation box (walls)
  double inv_L[3];
  double h[3];       // cell edge
  double inv_h[3];
  int nc[3];         // cells per dimension
  int ncell;
  int cap;           // slots per cell
  int npbc;
  int n, n_mols;
  double beta;
  int use_pair, use_ewald, use_bond, use_ext;
  int n_types;
  double rc;         // real-space cutoff (reference's derived value)
  double rs, rs2;    // search radius = max(rc, largest LJ cutoff)
  double sqrt_alpha, lB;
  double self_coef;  // -lB sqrt(alpha/kPi)
  double dip_coef;   // lB 2 kPi / V_e
  int real_cell[3];  // reference image range for explicit pair sums
  int nkx, nky, nkz; // k-space extents (|l| <= n)
  int tab_w;         // phase-table width per (bead, old/new)
  int n_kh;          // half-space k vectors
  double k_bond, r0;
  double wall_cut;
  double wall_r3_k213;  // pow(1/k213, 3), host-evaluated
  double wall_r3_cut;   // pow(1/wall_cut, 3), host-evaluated
  int dipole, dipole_exact;
  double fp64_below;    // FP32 path: pairs closer than this use FP64 math
  int single_image;     // only the minimum image can lie within rc (real_pair)
  const double2* erfc_tab;  // erfc nodes (erfc_nodes), kErfcNodes entries
  const float2* erfc_tab32;  // the same nodes rounded to FP32 (erfc_nodes32)
};

// FP64 erfc(x), x >= 0, by the Taylor series about the nearest tabulated
// node x_i = i / kErfcPerUnit, with the derivatives of erfc in terms of
// Hermite polynomials (Abramowitz & Stegun 7.1.19: the (n+1)-th derivative of
// erfc(x) is -(2 / sqrt(pi)) (-1)^n H_n(x) e^{-x^2}):
//   erfc(x_i + t) = erfc(x_i) - E_i sum_{n >= 0} (-1)^n H_n(x_i) t^{n+1} / (n+1)!,
//   E_i = (2 / sqrt(pi)) e^{-x_i^2},   |t| <= 1 / (2 kErfcPerUnit) = 1/256.
// Node i holds {erfc(x_i), E_i}, each rounded once from extended precision
// on the host.  Term n is about (2 x_i |t|)^n / n! of the first, and 2 x_i |t|
// <= 1/16 below x = 8, so the eight terms kept leave a relative remainder
// below (1/16)^9 / 9! < 5e-17; with the rounding of the nodes and of the sum
// the result is within a few ulp of erfc(x).  Beyond kErfcMax erfc(x) <
// 1.2e-29 and is taken as 0.  It keeps few registers live (no branch over
// argument ranges, no exp), which the pair kernels need.
constexpr int kErfcPerUnit = 128;
constexpr double kErfcMax = 8.0;
constexpr int kErfcNodes = 8 * kErfcPerUnit + 1;

// Node of x for erfc_nodes (0 where the node is not used: x at or beyond
// kErfcMax, or not a number), so that a kernel can issue the node's load
// before it needs the value.
__device__ __forceinline__ int erfc_node_index(double x) {
  return x < kErfcMax ? max(0, (int)(x * kErfcPerUnit + 0.5)) : 0;
}

// erfc(x) from its node i = erfc_node_index(x) and the node's table entry.
__device__ __forceinline__ double erfc_from_node(double x, int i, double2 node) {
  if (!(x < kErfcMax)) return 0.0;
  const double xi = (double)i * (1.0 / kErfcPerUnit);
  const double t = x - xi;   // exact: x and x_i are within a factor of 2 (or x_i = 0)
  // q_n = (-1)^n H_n(x_i): q_0 = 1, q_1 = -2 x_i, q_{n+1} = -2 x_i q_n - 2 n q_{n-1};
  // w_n = t^{n+1} / (n+1)!
  const double m2x = -2.0 * xi;
  double qm = 1.0, q = m2x;
  double w = t;
  double s = w;
  w *= t * 0.5;
  s += q * w;
#pragma unroll
  for (int n = 1; n < 7; n++) {
    const double qn = m2x * q - (2.0 * n) * qm;
    qm = q;
    q = qn;
    w *= t * (1.0 / (n + 2));
    s += q * w;
  }
  return node.x - node.y * s;
}

__device__ __forceinline__ double erfc_nodes(const Geo& g, double x) {
  const int i = erfc_node_index(x);
  return erfc_from_node(x, i, g.erfc_tab[i]);
}

// FP32 erfc(x) of the FP32 pair path: the same series about the same nodes
// (FP32-rounded), four terms: the remainder, about (2 x_i |t|)^4 / 5! of the
// first term, is below 1/16^4 / 120 < 1.3e-7 of it, and the first term E_i t
// is at most 2 x_i |t| <= 1/16 of erfc(x_i) below x = 8, so the result
// carries the FP32 rounding of the nodes and of the sum only (a few ulp).
__device__ __forceinline__ int erfc_node_index32(float x) {
  return x < (float)kErfcMax ? max(0, (int)(x * kErfcPerUnit + 0.5f)) : 0;
}

__device__ __forceinline__ float erfc_from_node32(float x, int i, float2 node) {
  if (!(x < (float)kErfcMax)) return 0.f;
  const float xi = (float)i * (1.0f / kErfcPerUnit);
  const float t = x - xi;
  const float m2x = -2.0f * xi;
  // q_0 = 1, q_1 = -2 x_i, q_2 = 4 x_i^2 - 2, q_3 = -2 x_i q_2 - 4 q_1
  const float q2 = m2x * m2x - 2.0f, q3 = m2x * q2 - 4.0f * m2x;
  const float t2 = t * t;
  const float s = t + t2 * (0.5f * m2x + t * ((1.0f / 6.0f) * q2 + t * ((1.0f / 24.0f) * q3)));
  return node.x - node.y * s;
}

__device__ __forceinline__ float erfc_nodes32(const Geo& g, float x) {
  const int i = erfc_node_index32(x);
  return erfc_from_node32(x, i, g.erfc_tab32[i]);
}

// LJ pair table entry for (ti, tj): sigma, epsilon, cutoff distance (strict
// r < cut), energy shift; computed on the host with the reference's own
// expressions (potential_truncated_lj.cc:49-85).
struct LJPair {
  double sigma, eps, cut, shift;
};

// FP32 copy for the reduced-precision pair path: sigma^2, 4 eps, cut^2, shift.
struct LJPair32 {
  float sigma2, eps4, shift;
  double cut;
};

struct WallType {
  double sigma, eps;
  int kind;          // 0 plain, 1 L, 2 R
  int pad;
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ double wrap_coord(double x, double L, double invL) {
  double w = x - L * floor(x * invL);
  if (w >= L) w -= L;
  if (w < 0) w += L;
  return w;
}

__device__ __forceinline__ int cell_coord(double xw, double inv_h, int n) {
  int c = (int)floor(xw * inv_h);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

__device__ __forceinline__ int floordiv(int a, int n) {
  return a >= 0 ? a / n : -((-a + n - 1) / n);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// Deterministic block reduction (fixed lane/wave order).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / kWave; w++) r += sh[w];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// Fixed-order wave reduction (shuffle down).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;  // valid in lane 0
}

// Orders one wave's LDS writes before its lanes read each other's entries.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The reference's minimum-image difference vector b - a
// (GetDistVector, misc.cc:43-56).
__device__ __forceinline__ void min_image(const double* a, const double* b,
                                          const double* L, int npbc, double* d) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double di = b[i] - a[i];
    if (i < npbc) di -= L[i] * round(di / L[i]);
    d[i] = di;
  }
}

// Bead::BBDist (bead.cc:162-176).
__device__ __forceinline__ double bb_dist(const double* a, const double* b,
                                          const double* L, int npbc) {
  double s = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double di = a[i] - b[i];
    if (i < npbc) {
      di -= L[i] * round(di / L[i]);
      if (fabs(di) > 0.5 * L[i]) di = L[i] - fabs(di);
    }
    s += di * di;
  }
  return sqrt(s);
}

// Slab dipole energy 2 pi l_B M_z^2 / V (PotentialEwaldCoul::DipoleE,
// potential_ewald_coul.cc:415-428) with coef = 2 pi l_B / V.  Each product is
// rounded on its own (no fused multiply-add with a later subtraction), so an
// unchanged M_z gives the same energy bit for bit, as the reference's
// repeated evaluation does, and the lagged dE of such a move is exactly 0.
__device__ __forceinline__ double dipole_energy(double coef, double mz) {
#pragma clang fp contract(off)
  return coef * mz * mz;
}

// e_new - e_old of two dipole energies, not fused with their products.
__device__ __forceinline__ double dipole_change(double e_new, double e_old) {
#pragma clang fp contract(off)
  return e_new - e_old;
}

// PotentialEwaldCoul::DipoleEDiff (potential_ewald_coul.cc:473-530): the
// slab dipole energy change coef (M_n^2 - M_o^2) when a bead (q z = a) and
// its counterion (b; 0 for a neutral chain) join a system of dipole M_o, as
// the reference evaluates it.
__device__ __forceinline__ double dipole_ediff(double coef, double mo, double a, double b) {
#pragma clang fp contract(off)
  double mn = mo + a;
  mn += b;
  return coef * (mn * mn - mo * mo);
}

// Truncated-shifted LJ / WCA at distance r (potential_truncated_lj.cc:49-85).
__device__ __forceinline__ double lj_energy(double r, const LJPair& p) {
  if (r <= 0) return kVeryLargeEnergy;
  if (r < p.cut) {
    double s = p.sigma / r;
    double s2 = s * s;
    double r6 = s2 * s2 * s2;
    return 4 * p.eps * (r6 * r6 - r6) - p.shift;
  }
  return 0.0;
}

// Real-space pair sum over the reference's image range
// (PotentialEwaldCoul::PairEnergyReal, potential_ewald_coul.cc:134-164).
__device__ inline double real_pair_ref(const double* a, double qa, const double* b,
                                       double qb, const Geo& g) {
  if (qa * qb == 0) return 0.0;
  double d[3];
  min_image(a, b, g.L, g.npbc, d);
  double pref = g.lB * qa * qb;
  double e = 0;
  for (int i = -g.real_cell[0]; i <= g.real_cell[0]; i++)
    for (int j = -g.real_cell[1]; j <= g.real_cell[1]; j++)
      for (int k = -g.real_cell[2]; k <= g.real_cell[2]; k++) {
        double x = d[0] + i * g.L[0], y = d[1] + j * g.L[1], z = d[2] + k * g.L[2];
        double r = sqrt(x * x + y * y + z * z);
        if (r > 0 && r <= g.rc) e += pref * erfc_nodes(g, g.sqrt_alpha * r) / r;
      }
  return e;
}

// The same pair sum when only the minimum image can lie within rc: every
// periodic edge is longer than 2 rc (g.single_image), and along a
// non-periodic z the separation is shorter than the padded period minus rc.
// It is then the (0, 0, 0) term of the image loop above, equal to that loop
// bit for bit since every other image falls beyond rc.
__device__ __forceinline__ double real_pair(const double* a, double qa, const double* b, double qb,
                                           const Geo& g) {
  if (qa * qb == 0) return 0.0;
  double d[3];
  min_image(a, b, g.L, g.npbc, d);
  if (!g.single_image || (g.npbc < 3 && !(fabs(d[2]) + g.rc * (1 + 1e-9) < g.L[2])))
    return real_pair_ref(a, qa, b, qb, g);
  const double pref = g.lB * qa * qb;
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double e = 0;
  if (r > 0 && r <= g.rc) e += pref * erfc_nodes(g, g.sqrt_alpha * r) / r;
  return e;
}

// PotentialTruncatedLJWall::BeadEnergy (potential_truncated_lj_wall.cc:46-134)
__device__ inline double wall_energy(double z, const WallType& w, const Geo& g) {
  const double Lz = g.Lsim[2];
  double energy = 0;
  const double sigma = w.sigma, epsilon = w.eps;
  const double R0 = 3 * k213 * sigma;
  const double K = 1;
  const double A = 2.59807621135;
  if (epsilon == 0) return 0.0;
  if (z <= 0 || z >= Lz) return kVeryLargeEnergy;
  if (g.wall_cut < 0 && w.kind == 0) {
    double r3_ref = g.wall_r3_k213;
    double energy_ref = A * epsilon * (r3_ref * r3_ref - r3_ref);
    if (z < k213 * sigma) {
      double r3 = sigma / z; r3 = r3 * r3 * r3;
      energy += A * epsilon * (r3 * r3 - r3) - energy_ref;
    }
    if (Lz - z < k213 * sigma) {
      double r3 = sigma / (Lz - z); r3 = r3 * r3 * r3;
      energy += A * epsilon * (r3 * r3 - r3) - energy_ref;
    }
  } else if (w.kind == 1) {
    double r3 = sigma / z; r3 = r3 * r3 * r3;
    energy += A * epsilon * (r3 * r3 - r3);
    double u = z / R0;
    energy += -0.5 * K * R0 * R0 * log(1 - u * u);
    if (Lz - z < k213 * sigma) {
      double r3b = sigma / (Lz - z); r3b = r3b * r3b * r3b;
      energy += A * epsilon * (r3b * r3b - r3b);
    } else {
      double r3b = g.wall_r3_k213;
      energy += A * epsilon * (r3b * r3b - r3b);
    }
  } else if (w.kind == 2) {
    double r3 = sigma / (Lz - z); r3 = r3 * r3 * r3;
    energy += A * epsilon * (r3 * r3 - r3);
    double u = (Lz - z) / R0;
    energy += -0.5 * K * R0 * R0 * log(1 - u * u);
    if (z < k213 * sigma) {
      double r3b = sigma / z; r3b = r3b * r3b * r3b;
      energy += A * epsilon * (r3b * r3b - r3b);
    } else {
      double r3b = g.wall_r3_k213;
      energy += A * epsilon * (r3b * r3b - r3b);
    }
  } else {
    double r3_ref = g.wall_r3_cut;
    double energy_ref = A * epsilon * (r3_ref * r3_ref - r3_ref);
    if (z < g.wall_cut) {
      double r3 = sigma / z; r3 = r3 * r3 * r3;
      energy += A * epsilon * (r3 * r3 - r3) - energy_ref;
    }
    if (Lz - z < g.wall_cut) {
      double r3 = sigma / (Lz - z); r3 = r3 * r3 * r3;
      energy += A * epsilon * (r3 * r3 - r3) - energy_ref;
    }
  }
  return energy;
}

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11): counter-based generator, 10 rounds, Weyl key schedule.
struct Philox {
  static __device__ __forceinline__ uint4 round1(uint4 c, uint2 k) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    return make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
  }
  static __device__ __forceinline__ uint4 gen(uint4 c, uint2 k) {
#pragma unroll
    for (int i = 0; i < 10; i++) {
      c = round1(c, k);
      k.x += 0x9E3779B9u;
      k.y += 0xBB67AE85u;
    }
    return c;
  }
};

// Uniform double in [0, 1] from 32 random bits, with the same granularity as
// the reference's (double)mt()/mt.max().
__device__ __forceinline__ double u01_32(uint32_t x) { return (double)x / 4294967295.0; }

}  // namespace plum
