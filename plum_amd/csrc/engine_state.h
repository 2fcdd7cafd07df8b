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
};

// Parameters of the device-resident loop, by value.
struct McDev {
  double cum_prob[5];     // move_prob as given (cumulated like the reference)
  double move_size, bond_len;
  int vary_bond, n_chains, n_ions;
  unsigned long long seed;
};

}  // namespace plum
