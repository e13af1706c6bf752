// Per-pattern numeric plan for the multifrontal LDL^T and the supernodal
// triangular solves: everything here depends only on the sparsity pattern,
// so it is built once on the host and reused for every set of values.
#pragma once

#include <cstdint>
#include <vector>

#include "host/sparse.hpp"

namespace mpir {

// Rows per panel-step task (at most 64).
constexpr int kPanelRows = 64;
// Row / column block of the big-front triangular solves.
constexpr int kSolveRB = 64;
// R rows per backward partial-sum task.
constexpr int kRpartRows = 512;
// Rows per forward task of a big front's inverse block (partial sums).
constexpr int kXcRows = 512;
// Tile edge of the trailing-update tasks.
constexpr int kTile = 64;
// Edge of the extend-add assembly tiles.
constexpr int kAsmTile = 64;

struct Task4 {
  int32_t a, b, c, d;
};

// Kinds of the dependency-driven solve tasks (Task4.a); b = first index into
// the kind's own task list, c = number of entries (up to kDagGroup for the
// one-wavefront kinds, else 1).
enum : int32_t { kDagSmall = 0, kDagGather, kDagXc, kDagU, kDagRpart, kDagBxc };
constexpr int kDagGroup = 4;
// Per-supernode entries of the solve dependency counts.
enum : int32_t { kMetaParent = 0, kMetaNChild, kMetaNGather, kMetaNXc, kMetaNU, kMetaNRpart, kMetaNBFinal,
                 kMetaStride = 8 };

struct FactorPlan {
  Supernodes sn;
  // children of each supernode, ascending
  std::vector<int32_t> child_ptr, child_list;
  // update-matrix pool: U_s is nrow_s x nrow_s column-major at upd_off[s]
  std::vector<int64_t> upd_off;
  int64_t upd_pool = 0;
  // per level: small fronts, big fronts
  std::vector<int32_t> small_ptr, small_list;
  std::vector<int32_t> big_ptr, big_list;
  // per level: assembly tasks (s, c0, c1, 0) over big fronts
  std::vector<int32_t> asm_ptr;
  std::vector<Task4> asm_task;
  // per assembly task, the children that land in it: (child, b0, b1, a0, a1)
  // = child columns [b0, b1) and child rows [a0, a1) mapping into the task
  std::vector<int32_t> asm_child_ptr;  // per task + 1
  std::vector<int32_t> asm_child;      // 5-tuples
  // per (level, step): panel tasks (s, row0, row1, step) and update tiles
  // (s, tile_row0, tile_col0, step); indexed through step_ptr. A front's
  // first panel task of a step has row0 = the step's first column: it owns
  // the diagonal block (writes L_kk, D, D^{-1}); the others cover at most
  // kPanelRows rows below it (or identity rows m ..)
  std::vector<int32_t> level_step_ptr;  // nlevels+1 into the step arrays
  std::vector<int32_t> pan_ptr;          // nsteps_total+1
  std::vector<Task4> pan_task;
  std::vector<int32_t> upd_ptr;          // nsteps_total+1
  std::vector<Task4> upd_task;
  // per step, the first panel / update task on the identity rows below R_s
  // (the inverse of the diagonal block); those tasks end the step's range
  std::vector<int32_t> pan_aug_ptr, upd_aug_ptr;  // nsteps_total
  // triangular solves of big fronts, per level (pairs (s, block)):
  //   sfwd: 64-row blocks of the front, gathering b and the children's
  //         update vectors; sxc: triples (s, column block, row chunk) of
  //         x_c = (L_11^{-T} D^{-1})^T v_c, partial sums over kXcRows rows at
  //         xfp_off[s] + (block * nchunk + chunk) * kSolveRB,
  //         nchunk = ceil(nc / kXcRows);
  //   su:   64-row blocks of R_s, u = v_R - L_21 D x_c;
  //   sbwd: row blocks, x_c = L_11^{-T} (x_c - L_21^T x_R)
  std::vector<int32_t> sfwd_ptr, sxc_ptr, su_ptr, sbwd_ptr;  // nlevels+1
  std::vector<int32_t> sfwd_task, sxc_task, su_task, sbwd_task;
  // per forward gather task, the children whose update vectors land in its
  // rows: (child, q0, q1) = child rows [q0, q1)
  std::vector<int32_t> sfwd_child_ptr;        // per forward task + 1
  std::vector<int32_t> sfwd_child;            // triples
  // backward contribution of the rows below the columns (R rows): tasks
  // (s, column block, row chunk) per level, partial sums at
  // part_off[s] + (block * nchunk + chunk) * kSolveRB
  std::vector<int32_t> rpart_ptr;             // nlevels+1
  std::vector<int32_t> rpart_task;            // triples
  std::vector<int64_t> part_off;              // per supernode
  int64_t npart = 0;
  std::vector<int64_t> xfp_off;               // per supernode
  int64_t nxfp = 0;
  // dependency-driven solves: every task above, forward in level order
  // (small groups, gathers, column blocks, R blocks per level) and backward
  // in reverse level order (R partials, column blocks, small groups); each
  // task depends only on tasks before it. dag_meta: kMetaStride entries per
  // supernode (parent, children, and the task counts the waits compare with)
  std::vector<Task4> dag_fwd, dag_bwd;
  std::vector<int32_t> dag_meta;
  // for each entry of the ORIGINAL upper CSC: destination in the panel pool
  std::vector<int64_t> val_dst;
  // statistics
  int64_t panel_entries = 0;
  int64_t nsmall = 0, nbig = 0;
};

// Schedule and storage layout of a supernodal partition (no value map).
FactorPlan build_plan_from_supernodes(Supernodes sn);

// perm: new -> old. C: permuted upper CSC with val_map (C entry -> original
// entry). parent/Lnz: etree of C.
FactorPlan build_factor_plan(const CscUpper& C, const std::vector<int32_t>& val_map,
                             const int32_t* parent, const int32_t* Lnz);
// Same with a given (relaxed) supernode partition.
FactorPlan build_factor_plan_forced(const CscUpper& C, const std::vector<int32_t>& val_map,
                                    const int32_t* parent, const std::vector<int32_t>& first);

// The host-side symbolic phase of the solver: ordering (0 natural, 1 AMD,
// 2 nested dissection; AMD/ND followed by an elimination-tree postorder),
// optional relaxed supernodes, the permuted upper CSC and the plan.
struct Analysis {
  std::vector<int32_t> perm;     // new -> old
  CscUpper C;                    // triu(P A P^T), pattern only
  std::vector<int32_t> val_map;  // C entry -> A entry
  std::vector<int32_t> parent, Lnz;
  int64_t nnz_l = 0;
  double flops = 0.0;            // sum Lnz_j (Lnz_j + 2) + n
  bool relaxed = false;
  FactorPlan plan;
};
Analysis analyze_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, bool relax);

}  // namespace mpir
