// Device-resident state of one system (layout in HBM) and the small records
// the host and the kernels exchange.
#pragma once
#include "engine_dev.h"

namespace plum {

// The pending trial move, written by the caller (host path) or by the
// device proposal kernel.
struct MoveRec {
  int n_moved;
  int mol;
  int move_type;   // 0 bead, 1 COM, 2 pivot, 3 crankshaft, 4 reptation, -1 external
  int valid;       // 0: no move attempted (e.g. no chain to move)
  int contig;      // moved beads are exactly the index range [lo, hi]
  int lo, hi;
  int pad;
};

constexpr int kMaxProp = 16;                                  // exactly evaluated proposals per batch
constexpr int kMaxScreen = 32;                                // screened proposals per batch
constexpr int kMaxPairs = kMaxProp * (kMaxProp - 1) / 2;       // exact-slot pairs
constexpr int kKVals = kMaxProp + kMaxPairs;                   // k-space sums per batch
constexpr int kDescr = 64;                                     // surrogate descriptor length

// Pair slot of exact slots p != q.
__host__ __device__ __forceinline__ int pair_index(int p, int q) {
  const int a = p < q ? p : q, b = p < q ? q : p;
  return b * (b - 1) / 2 + a;
}

// One proposal of a speculative batch (written by the proposal kernel).
struct PropInfo {
  double bond_trial;   // spring energy of the molecule at its trial coordinates
  double ew;           // wall dE of the moved beads
  double dmz;          // sum q (z' - z) of the moved beads
  double c[3], rad;    // bounding sphere of the moved beads, old and new
  int forbid;          // a moved bead entered a wall (1e8)
  int n, mol, type;    // moved beads, molecule, move type
  int valid, skipped;  // attempted move / dropped after a cell overflow
  int prev;            // earlier proposal of the same molecule in the batch, -1 none
  int len;             // beads of the molecule
  double bcur;         // spring energy of the molecule at its current coordinates
};

// Plan and reduced sums of a speculative batch.  Proposals live in screen
// slots (up to kMaxScreen); the ones evaluated exactly get exact slots (up
// to kMaxProp), which carry dS(k), the pair partial sums and the cross terms.
struct SpecPlan {
  int n_prop;                    // proposals made this batch
  int n_decide;                  // proposals that can be decided (stops before a pass with no exact slot)
  int n_exact;                   // exact slots in use
  int n_near;                    // exact-slot pairs within interaction range
  int exact_of[kMaxScreen];      // exact slot of a proposal, -1 if none
  int screen_of[kMaxProp];       // proposal of an exact slot
  int pass1[kMaxScreen];         // first stage of delayed acceptance passed (1 without it)
  int pair_p[kMaxPairs], pair_q[kMaxPairs];
  int pair_off[kMaxPairs + 1];   // cross-term blocks per near pair (prefix)
  int qwave;                     // waves per real-space query this batch (query_waves)
  int q_off[kMaxProp + 1];       // real-space query blocks per exact slot (prefix)
  int mm_off[kMaxProp + 1];      // moved-moved blocks per exact slot (prefix)
  unsigned acc_mask;             // exact slots accepted by the last resolve: S(k) += dS pending
  int pad;
  double de_s[kMaxScreen];       // surrogate dE of each proposal (0 without it)
  double ksum[kKVals];           // k space: dE per exact slot, then 2 sum c_k Re(dS_p* dS_q)
  double2 rsum[kMaxProp];        // (LJ, real) dE per exact slot against the unmoved state
  double2 xsum[kMaxPairs];       // (LJ, real) cross term per near pair
};

// One decided proposal of the device loop (bounded record, for replay
// checks): what the Metropolis test was given and what it decided.
struct DecRec {
  unsigned long long move;   // move number (Philox counter) of the proposal
  int mol, n, type, accepted;
  int lo, hi, contig, evaluated;   // evaluated: the exact dE was computed
  int early, pass, pad0, pad1;     // early: an early exit (1e8); pass: first stage passed
  double terms[8];           // per-term dE, cross terms of earlier acceptances included
  double total;              // dE with the reference's early exits
  double de_s;               // surrogate dE of the first stage (0 without it)
};

// Control block: scalars of the energy state and the run.
struct Ctl {
  double Mz;            // sum q z over current coordinates
  double dMz;           // sum q (z' - z) of the pending move
  double cur_dipl;      // the reference's current_dipl_E
  double trial_dipl;    // the reference's trial_dipl_E
  double bond_trial;    // spring energy of the moved molecule (trial)
  double terms[8];      // dE of the pending move, per term
  double total;         // dE total with the reference's early exits
  double running[8];    // running totals per term
  double full[8];       // last from-scratch energy
  int accept;           // decision of the pending move
  int overflow;         // a cell ran out of slots (cell list incomplete)
  int pending;          // a dE is waiting for commit/rollback
  int sk_pending;       // S(k) += dS(k) still to be applied
  unsigned long long move_counter;   // Philox counter of the device loop
  long long attempted[5];
  long long accepted[5];
  long long valid_moves;
  long long skipped;    // device-loop moves skipped after a cell overflow
  // delayed acceptance
  long long s1_rejected;   // first-stage rejections (no exact evaluation needed)
  long long s1_passed;     // first-stage passes (exact evaluation, second stage)
  long long train_evals;   // first-stage rejections evaluated for training
  long long stale_cuts;    // batches ended early by a stale surrogate value
};

struct Dev {
  // beads, index order
  double *x, *y, *z, *q;
  int *type, *mol_of, *mol_start;
  int *bcell, *bslot;     // cell and slot of each bead
  unsigned char* moved;   // bead is in the pending move
  int* mv_slot;           // index in the moved list, -1 otherwise
  // cell list: ccount = slots in use (high-water mark, tombstones included);
  // cbead = bead index or -1 (tombstone)
  int *ccount, *cbead;
  double4* cpos;          // wrapped x,y,z and q (FP64)
  float4* cpos32;         // x,y,z relative to the cell origin and q (FP32 path)
  int* ctype;
  // k space
  int4* kl;
  int2* krun;             // runs of k vectors: (first, count), one (l_x, l_y), consecutive l_z
  double* ck;
  double2 *Sk, *dS;
  double2* tab;           // phase tables
  // potential tables
  LJPair* lj;
  LJPair32* lj32;
  WallType* wall;
  double* bond_cur;       // per-molecule spring energy (current)
  // pending move
  MoveRec* rec;
  int* mv_idx;
  double* mv_xyz;         // caller's trial xyz (host path)
  double4 *mv_old, *mv_new;
  int* mv_type;
  Ctl* ctl;
  // partial sums
  double* part_k;
  double2* part_r;
  double2* part_mm;
  double2* sk_part;
  double* fpart;
  // device Monte Carlo
  int *chain_ids, *ion_ids;
  // speculative batches: proposal s's moved beads at s * maxm in mv_old /
  // mv_new / mv_type / smv_idx, its phase tables at s * maxm * 2 * tab_w;
  // exact slot e's dS(k) at e * n_kh
  MoveRec* recs;
  int* smv_idx;
  PropInfo* prop;
  SpecPlan* plan;
  double* part_ks;        // [kKVals][k blocks]
  double2* part_x;        // cross-term block partials
  // delayed acceptance (null without the surrogate): descriptor rows and
  // network outputs of every proposal (row (2 s + on) * maxm + m, on = 0 old
  // / 1 new), and the ring of training samples (rows, target, bead count)
  float* sur_x;
  float* sur_f;
  float* tr_x;
  double* tr_target;
  int* tr_n;
  int* tr_ctl;            // [0] samples written (ring head), [1] ring capacity
  // decision record (null when off): record k holds proposal metadata and
  // its maxm trial positions / bead indices at k * maxm; rec_ctl[0] counts
  // decided proposals since the record was armed, rec_ctl[1] is the capacity
  DecRec* drec;
  double4* drec_pos;
  int* drec_idx;
  int* drec_ctl;
};

// Parameters of the device-resident loop, by value.
struct McDev {
  double cum_prob[5];     // move_prob as given (cumulated like the reference)
  double move_size, bond_len;
  int vary_bond, n_chains, n_ions;
  unsigned long long seed;
  unsigned long long ctr_limit;   // proposals stop at this move number
};

}  // namespace plum
