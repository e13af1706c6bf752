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
// Rows per task of the column-block products of the solves (W^T v, L_21^T x)
// and of the factor's inverse post-pass.
constexpr int kXcRows = 256;
// Columns per task of the row-block products of the solves (L_21 y, W z).
constexpr int kGemvCols = 256;
// Tile edge of the trailing-update tasks.
constexpr int kTile = 64;
// Panel steps per update group (FactorTuning::upd_group): a step updates
// the trailing columns of its own group (rank 64); the columns right of the
// group are updated once, at the group's last step, with all its steps
// (rank upd_group * 64).
// Update task field d: the last step k, the number of steps ending at k
// that the task applies, and a flag restricting the task to panel columns
// (c < nc): d = k | (steps << kUpdStepsShift) | (panel_only << kUpdPanelBit).
constexpr int kUpdStepsShift = 20;
constexpr int kUpdPanelBit = 30;
// Edge of the extend-add assembly tiles.
constexpr int kAsmTile = 64;
// Narrow big fronts: at most kSolveRB columns and at most
// SolveTuning::narrow_rows rows below the diagonal block. Their forward
// solve is one task per kNarrowChunk rows of L_21 (each forms y = W^T v
// itself, so no task waits on another task of the same front) and their
// backward solve one task per front.
constexpr int kNarrowChunk = 256;

// Tunable shape of the triangular-solve schedule (it changes which task
// kinds cover a front, never the factor's layout).
struct SolveTuning {
  int32_t narrow_rows = 1024;  // largest nrow of a narrow front (0: none)
};
// Tunable shape of the factor's step schedule (the trailing-update tasks;
// the storage layout does not change).
struct FactorTuning {
  int32_t upd_group = 4;    // panel steps per update group, 1 .. 64
  int32_t super_tiles = 8;  // edge of the tile super-blocks, in tiles, 1 .. 64
};
inline bool same_tuning(const SolveTuning& a, const SolveTuning& b) { return a.narrow_rows == b.narrow_rows; }
inline bool same_tuning(const FactorTuning& a, const FactorTuning& b) {
  return a.upd_group == b.upd_group && a.super_tiles == b.super_tiles;
}

struct Task4 {
  int32_t a, b, c, d;
};

// Kinds of the dependency-driven solve tasks (Task4.a). Small fronts:
// (kDagSmall, first index into small_list, count <= kDagGroup, 0); tiles of
// a big front: (kind, s, block, chunk).
enum : int32_t { kDagSmall = 0, kDagFx, kDagFu, kDagBr, kDagBx, kDagFm, kDagBm };
constexpr int kDagGroup = 4;
// Per-supernode entries of the solve schedule: parent, children, column
// blocks, R blocks, backward R-chunk tasks, backward completions (column
// blocks, or 1 for a small front), column chunks, R chunks.
enum : int32_t { kMetaParent = 0, kMetaNChild, kMetaNFx, kMetaNFu, kMetaNBr, kMetaNBFinal, kMetaNH, kMetaNHR,
                 kMetaStride = 8 };

// Dependency-driven factorization: the big fronts of the levels at and
// above FactorPlan::fdag_level in one persistent launch. Task kinds:
//   asm   one extend-add tile of a front (x = index of the assembly task)
//   diag  the diagonal block of panel step k of a front (x = k): L_kk, D,
//         D^{-1} and Z_kk = L_kk^{-T} D_k^{-1} in the block's identity rows
//   pan   rows [x, y) of panel step z: L = X Z_kk (rows below the block, or
//         identity rows of earlier blocks)
//   upd   a trailing-update tile (x = row, y = column, z = step code as in
//         the per-level update tasks)
// A task waits until counter wa >= ta and, for wb, wc >= 0, until counter
// wb >= tb and counter wc >= tc.
// Counters (index from fdag_meta): per front asm tasks done, DAG children
// complete, update tasks done; per (front, step) diagonal blocks done and
// panel tasks done (the diagonal block counts as one of them); per (front,
// 64-column block) update tasks done of the tiles that start on the block.
// Update tiles at a front's last step start at column nc, which is not on a
// block boundary when 64 does not divide nc; they touch only the update
// matrix, wait on the two blocks they overlap, and count only towards the
// front's completion.
enum : int32_t { kFAsm = 0, kFDiag = 1, kFPan = 2, kFUpd = 3 };
constexpr int kFKindShift = 28;
struct FTask {
  int32_t sk;  // supernode | kind << kFKindShift
  int32_t x, y, z;
  int32_t wa, ta, wb, tb, wc, tc;
};
// Per supernode in fdag_meta: first step counter, first column-block
// counter, update tasks, parent (or -1 when the parent is not in the launch).
enum : int32_t { kFMetaStep = 0, kFMetaCb, kFMetaNUpd, kFMetaParent, kFMetaStride = 4 };

struct FactorPlan {
  Supernodes sn;
  // children of each supernode, ascending
  std::vector<int32_t> child_ptr, child_list;
  // update-matrix pool: U_s is nrow_s x nrow_s column-major at upd_off[s];
  // dag_pool: no block is reused from fdag_level on (the layout the
  // dependency-driven factorization needs; larger)
  std::vector<int64_t> upd_off;
  int64_t upd_pool = 0;
  bool dag_pool = false;
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
  // (s, tile_row0, tile_col0, step | steps << kUpdStepsShift); indexed
  // through step_ptr. A front's
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
  // Triangular solves of big fronts (W = L_11^{-T} below the panel), task
  // triples per level:
  //   fx (s, column block j, row chunk h): partial sums of y = W^T v over
  //      kXcRows rows of 64 columns, v gathered on those rows
  //   fu (s, R block b, column chunk h): partial sums of L_21 y over
  //      kGemvCols columns for 64 R rows; chunk 0 also stores v_R, so that
  //      u = v_R - (partial sums) is formed by whoever gathers it
  //   br (s, column block j, R chunk h): partial sums of L_21^T x_R over
  //      kXcRows R rows
  //   bx (s, row block i, column chunk h): partial sums of W z,
  //      z = D^{-1} y - L_21^T x_R, over kGemvCols columns (from column
  //      64 i on) for 64 rows; the last task of a block stores x
  // y, L_21 y and L_21^T x_R stay as partial sums (per 64-entry block and
  // chunk, at ypart_off / upart_off / rpart_off of the front) that their
  // readers add in a fixed order.
  std::vector<int32_t> fx_ptr, fu_ptr, br_ptr, bx_ptr;      // nlevels+1, in tasks
  std::vector<int32_t> fx_task, fu_task, br_task, bx_task;  // triples
  // narrow fronts (kNarrowChunk): fm (s, R chunk h, 0) gathers v, forms
  // y = W^T v and u = v_R - L_21 y on kNarrowChunk rows of R (the h = 0 task
  // stores y); bm (s, 0, 0) forms x = W (D^{-1} y - L_21^T x_R). Their u is
  // final (no partial sums) and their y has one partial sum per column.
  std::vector<int32_t> fm_ptr, bm_ptr;    // nlevels+1, in tasks
  std::vector<int32_t> fm_task, bm_task;  // triples
  // the factor's W = Z D pass: (s, column block j, row chunk h) over the
  // rows at or above the block's last column, every big front
  std::vector<int32_t> inv_task;  // triples
  SolveTuning tuning;
  FactorTuning ftuning;
  // row gathers of every front: the contributions to row r of front s are
  // entries [grow_ptr[frow_off[s] + r], grow_ptr[frow_off[s] + r + 1]) of
  // gent, in child order: (update-vector entry, first partial sum of L_21 y
  // for that child row or 0, number of those partial sums, 0)
  std::vector<int32_t> frow_off, grow_ptr;
  std::vector<Task4> gent;
  // per supernode: first column block in the bx block counters; offsets of
  // the partial-sum areas (entries)
  std::vector<int32_t> cb_base;
  std::vector<int64_t> ypart_off, upart_off, rpart_off, xpart_off;
  int64_t nypart = 0, nupart = 0, nrpart = 0, nxpart = 0;
  int32_t ncb_total = 0;
  // dependency-driven solves: every task above, forward in level order
  // (small groups, fx, fu per level) and backward in reverse level order
  // (br, bx, small groups); each task depends only on tasks before it.
  // dag_meta: kMetaStride entries per supernode
  std::vector<Task4> dag_fwd, dag_bwd;
  std::vector<int32_t> dag_meta;
  // dependency-driven factorization of the levels >= fdag_level (no small
  // fronts there; nlevels when there is none): task list in a topological
  // order, per-supernode counter offsets, counter layout
  int32_t fdag_level = 0;
  std::vector<FTask> fdag;
  std::vector<int32_t> fdag_meta;  // kFMetaStride per supernode
  int32_t fdag_nsteps = 0, fdag_ncb = 0;
  // for each entry of the ORIGINAL upper CSC: destination in the panel pool
  std::vector<int64_t> val_dst;
  // statistics
  int64_t panel_entries = 0;
  int64_t nsmall = 0, nbig = 0;
};

// Schedule and storage layout of a supernodal partition (no value map).
FactorPlan build_plan_from_supernodes(Supernodes sn, const FactorTuning& ft = {}, const SolveTuning& st = {});
// The factor's panel and update tasks per (level, step) of P's big fronts
// for tuning t (level_step_ptr, pan_*, upd_*).
void build_factor_steps(FactorPlan& P, const FactorTuning& t);
// dag_fwd, dag_bwd and dag_meta from the per-level task lists.
void build_dag_schedule(FactorPlan& P);
// First level of the dependency-driven factorization: one above the highest
// level that holds a small front.
int32_t factor_dag_level(const FactorPlan& P);
// upd_off, upd_pool and dag_pool of P's supernodes: blocks are released when
// their parent's level has been assembled, below factor_dag_level only when
// `dag` (the dependency-driven factorization runs fronts of several levels
// at once); with `dag` also the dependency-driven task list (fdag),
// without it that list is cleared. Sets fdag_level either way.
void build_update_pool(FactorPlan& P, bool dag);
// fdag, fdag_meta and the counter layout from the per-level factor lists
// (assembly, panel and update tasks); build_update_pool(P, true) calls it.
void build_factor_dag(FactorPlan& P);
// Counter indices of the dependency-driven factorization (constexpr: the
// kernels use the same functions).
struct FactorDagIndex {
  int32_t ns, nsteps;
  constexpr int32_t asm_done(int32_t s) const { return s; }
  constexpr int32_t child_done(int32_t s) const { return ns + s; }
  constexpr int32_t upd_done(int32_t s) const { return 2 * ns + s; }
  constexpr int32_t diag_done(int32_t step) const { return 3 * ns + step; }
  constexpr int32_t pan_done(int32_t step) const { return 3 * ns + nsteps + step; }
  constexpr int32_t col_done(int32_t cb) const { return 3 * ns + 2 * nsteps + cb; }
};
// The triangular-solve schedule of P's supernodes (task lists, partial-sum
// offsets, row gathers, dependency schedule) for tuning t; P.inv_task too.
void build_solve_schedule(FactorPlan& P, const SolveTuning& t);

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
// relax: 0 fundamental supernodes, 1 .. 3 relaxed (relax_supernodes level).
// The default, 3, merges fronts up to 64 columns with at most 80 % zeros and
// up to 128 with at most 20 %: on the 64^3 KKT system the tree has 18 levels
// instead of 25 (level 1), for 10 % more stored entries.
constexpr int kDefaultAmalgamation = 3;
Analysis analyze_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, int ordering, int relax);

}  // namespace mpir
