// Launch interfaces of the HIP kernels (gfx950).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../plan.hpp"

namespace mpir {

// Device view of a factorization: pattern arrays of the plan plus the
// numeric storage in factor precision T.
template <class T>
struct FactorArgs {
  const int32_t* first;      // ns+1
  const int32_t* nrow;       // ns
  const int64_t* rows_off;   // ns+1
  const int32_t* rows;       // sum nrow
  const int32_t* relidx;     // sum nrow
  const int64_t* panel_off;  // ns+1
  const int64_t* upd_off;    // ns
  const int32_t* child_ptr;  // ns+1
  const int32_t* child_list;
  T* panel;
  T* upool;
  T* D;
  T* Dinv;
  int* err;
};

// Device copies of the plan's task lists, with the host-side offsets the
// launch loop needs.
struct DeviceFactorPlan {
  int nlevels = 0;
  int64_t panel_entries = 0;
  const int32_t* small_list = nullptr;
  const int32_t* big_list = nullptr;
  const Task4* asm_task = nullptr;
  const Task4* pan_task = nullptr;
  const Task4* upd_task = nullptr;
  const int32_t* asm_child_ptr = nullptr;
  const int32_t* asm_child = nullptr;
  const int64_t* val_dst = nullptr;
  std::vector<int32_t> h_small_ptr, h_big_ptr, h_asm_ptr, h_level_step_ptr, h_pan_ptr, h_upd_ptr;
  std::vector<int32_t> h_pan_aug_ptr, h_upd_aug_ptr;
  // (s, column block, row chunk) tasks over the inverse rows of big fronts
  const int32_t* inv_task = nullptr;
  int n_inv_task = 0;
};

// rounding: storage format of the factor's working values (0 as T, 1 FP16,
// 2 BF16; only for T = float). Leaves W = L_11^{-T} below big panels.
template <class T>
void factor_numeric(const FactorArgs<T>& a, const DeviceFactorPlan& dp, const double* ax_scaled, int64_t nnz,
                    int rounding, hipStream_t st);
// Recomputes only the rows W = L_11^{-T} below each big front's panel from a
// factor already in place (L in the panel, D and Dinv set).
template <class T>
void invert_diag_blocks(const FactorArgs<T>& a, const DeviceFactorPlan& dp, hipStream_t st);
// The whole panel pool rounded to a 16-bit factor format; *nonfinite
// receives the number of entries outside the format's range.
template <class H>
void round_panel(int64_t n, const float* src, H* dst, int* nonfinite, hipStream_t st);
// y = 1 / x elementwise
template <class T>
void reciprocal(int n, const T* x, T* y, hipStream_t st);
template <class T>
void count_positive(int n, const T* D, int* out, hipStream_t st);

// ---------------------------------------------------------------- solves
// Device view of the factor for the triangular solves (factor storage TF,
// D^{-1} in TD, arithmetic in TX).
template <class TF, class TD>
struct SolveArgs {
  const int32_t* first;
  const int32_t* nrow;
  const int64_t* rows_off;
  const int32_t* rows;
  const int32_t* relidx;
  const int64_t* panel_off;
  const int32_t* child_ptr;
  const int32_t* child_list;
  const TF* panel;  // L, and W = L_11^{-T} below the panels of big fronts
  const TD* Dinv;
};

// Schedules of the triangular solves: one launch per task kind and level,
// or one persistent launch per direction driven by dependency counters.
constexpr int kScheduleLevels = 0;
constexpr int kScheduleDag = 1;

// Dependency-driven schedule: task lists in topological order (kDag* kinds)
// and per-front counts; counters: kNumSolveCounters * ns completion counts
// followed by kDispensers task dispensers per direction, each on its own
// 128-byte line; all zeroed before every solve.
constexpr int kNumSolveCounters = 6;
constexpr int kDispensers = 8;
constexpr int kDispenserStride = 32;
constexpr int dispenser_base(int ns) { return (kNumSolveCounters * ns + kDispenserStride - 1) / kDispenserStride * kDispenserStride; }
constexpr int dag_counter_count(int ns) { return dispenser_base(ns) + 2 * kDispensers * kDispenserStride; }
struct DagArgs {
  int ns = 0, nfwd = 0, nbwd = 0;
  const Task4* fwd = nullptr;
  const Task4* bwd = nullptr;
  const int32_t* meta = nullptr;  // kMetaStride per supernode
  const int32_t* small_list = nullptr;
  const int32_t* gather_task = nullptr;
  const int32_t* gather_child_ptr = nullptr;
  const int32_t* gather_child = nullptr;
  const int32_t* xc_task = nullptr;
  const int32_t* u_task = nullptr;
  const int32_t* rpart_task = nullptr;
  const int32_t* bwd_task = nullptr;
  const int64_t* xfp_off = nullptr;
  const int64_t* part_off = nullptr;
  int* counters = nullptr;
};

struct DeviceSolvePlan {
  int schedule = kScheduleDag;
  DagArgs dag;
  int* err = nullptr;  // bit 1: a dependency wait expired
  int nlevels = 0;
  const int32_t* small_list = nullptr;    // small fronts, grouped by level
  const int32_t* gather_task = nullptr;   // (s, row block) pairs, grouped by level
  const int32_t* gather_child_ptr = nullptr;  // per gather task + 1
  const int32_t* gather_child = nullptr;      // (child, q0, q1) triples
  const int32_t* xc_task = nullptr;       // (s, column block, row chunk) triples
  const int32_t* u_task = nullptr;        // (s, R row block) pairs
  const int32_t* bwd_task = nullptr;      // (s, row block) pairs
  const int32_t* rpart_task = nullptr;    // (s, column block, row chunk) triples
  const int64_t* part_off = nullptr;      // per supernode
  void* part = nullptr;                   // partial sums of the solve precision
  const int64_t* xfp_off = nullptr;       // per supernode
  void* xfp = nullptr;                    // forward partial sums of big fronts
  std::vector<int32_t> h_small_ptr, h_gather_ptr, h_xc_ptr, h_u_ptr, h_bwd_ptr, h_rpart_ptr;
};

// x (length n, permuted order) is overwritten by (L D L^T)^{-1} x.
// uvec: scratch of sum(nrow) entries of TX.
template <class TF, class TD, class TX>
void ldl_solve(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st);

// ---------------------------------------------------------------- vectors / SpMV
struct CsrDev {
  int32_t n = 0;
  const int32_t* p = nullptr;
  const int32_t* j = nullptr;
  const double* x = nullptr;
};

// y = A x (FP64, rows summed left to right, no fused multiply-add)
void spmv_f64(const CsrDev& A, const double* x, double* y, hipStream_t st);
// r = b - A x
void residual_f64(const CsrDev& A, const double* x, const double* b, double* r, hipStream_t st);

}  // namespace mpir
