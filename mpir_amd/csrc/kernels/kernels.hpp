// Launch interfaces of the HIP kernels (gfx950).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../plan.hpp"

namespace mpir {

// value of v in lane j of the wavefront (j uniform across the wave)
template <class T>
__device__ __forceinline__ T lane_value(T v, int j) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(b), j);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(b >> 32), j);
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), j));
  }
}

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
  // split panel steps: Z_kk of each front of a level (kPanelNB x kPanelNB,
  // column-major), slot zslot[s] (the front's place in its level's list)
  T* zbuf = nullptr;
  const int32_t* zslot = nullptr;
  int zcap = 0;  // slots in zbuf
};

// Device view of the dependency-driven factorization (plan.hpp, FTask):
// the task list, the per-supernode counter offsets, the assembly lists it
// indexes, and the counters (the task counters, then the dispenser lines).
struct FactorDagDev {
  const FTask* task = nullptr;
  int ntask = 0;
  const int32_t* meta = nullptr;  // kFMetaStride per supernode
  const Task4* asm_task = nullptr;
  const int32_t* asm_child_ptr = nullptr;
  const int32_t* asm_child = nullptr;
  int* counters = nullptr;
  int ns = 0, nsteps = 0, cnt_disp = 0, ncounters = 0;
  // when set: per task (its kind and supernode, the device wall clock when
  // it was drawn, when its waits ended, when it was counted, the workgroup)
  unsigned long long* trace = nullptr;
};

// Panel-step launches of at least this many tasks are split (the fronts'
// diagonal blocks eliminated once, DeviceFactorPlan::pan_split_min), when
// the Z_kk blocks of the level fit in this many bytes.
constexpr int kPanelSplitMin = 512;
constexpr int64_t kZbufMaxBytes = int64_t(1) << 30;

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
  // panel-step launches of at least pan_split_min tasks (0: none) run as
  // two launches, the fronts' diagonal blocks first (their Z_kk to
  // FactorArgs::zbuf), then the other rows with Z_kk read back, when the
  // level has at most zbuf_fronts big fronts
  int pan_split_min = 0, zbuf_fronts = 0;
  // per step, the tasks that own a diagonal block and all others, for the
  // two launches of a split step (offsets per step into each list)
  const Task4* pan_own = nullptr;
  const Task4* pan_rest = nullptr;
  std::vector<int32_t> h_pan_own_ptr, h_pan_rest_ptr;
  // levels >= dag_level in one dependency-driven launch (when use_dag; the
  // update-matrix pool must then be laid out for it: dag_pool)
  bool use_dag = false, dag_pool = false;
  int dag_level = 0;
  FactorDagDev dag;
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
void count_positive(int n, const T* D, const T* Dinv, int* out, int* nonfinite, hipStream_t st);

// ---------------------------------------------------------------- solves
// Entries allocated past the end of a panel pool: the solves' 16-byte loads
// may run up to 7 entries past the last entry of a tile.
constexpr int64_t kPanelSlack = 64;

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

// Counters of the solves (int), zeroed before every application:
//   kNumSolveCounters per supernode (completion counts), then one per
//   column block of the big fronts (bx partial sums in), then per direction
//   kDispensers task dispensers and one arrival counter, each on its own
//   128-byte line.
constexpr int kNumSolveCounters = 5;
constexpr int kDispensers = 16;
constexpr int kDispenserStride = 32;
constexpr int kDispenserLines = kDispensers + 1;

// Device view of the solve plan (task lists, gather table, offsets, work
// areas of the solve precision).
struct SolveDev {
  int ns = 0, nfwd = 0, nbwd = 0;
  const Task4* fwd = nullptr;  // dependency-driven task lists
  const Task4* bwd = nullptr;
  const int32_t* meta = nullptr;  // kMetaStride per supernode
  const int32_t* small_list = nullptr;
  const int32_t* fx_task = nullptr;
  const int32_t* fu_task = nullptr;
  const int32_t* br_task = nullptr;
  const int32_t* bx_task = nullptr;
  const int32_t* fm_task = nullptr;  // narrow fronts
  const int32_t* bm_task = nullptr;
  const int32_t* frow_off = nullptr;  // row gathers
  const int32_t* grow_ptr = nullptr;
  const Task4* gent = nullptr;
  const int32_t* cb_base = nullptr;
  const int64_t* ypart_off = nullptr;
  const int64_t* upart_off = nullptr;
  const int64_t* rpart_off = nullptr;
  const int64_t* xpart_off = nullptr;
  void* ypart = nullptr;  // partial sums (solve precision)
  void* upart = nullptr;
  void* rpart = nullptr;
  void* xpart = nullptr;
  int* counters = nullptr;
  int cnt_bx = 0, cnt_disp = 0, ncounters = 0;
  int* err = nullptr;  // bit 1: a dependency wait expired
  // entries between right-hand side r and r + 1: in x, in uvec, and in the
  // partial-sum areas (ldl_solve_multi)
  int64_t sx = 0, su = 0, syp = 0, sup = 0, srp = 0, sxp = 0;
  // when set, every kernel of an application returns at once if *gate == 0
  // (an iteration that has already stopped; read at kernel start, written
  // by earlier work of the same stream)
  const int* gate = nullptr;
  // when set (dependency-driven schedule): per task of the forward list, then
  // of the backward list, its kind | first supernode << 4 and the device wall
  // clock when it was drawn, when its wait ended, when it was done (<< 16 |
  // the workgroup)
  unsigned long long* trace = nullptr;
};

struct DeviceSolvePlan {
  int schedule = kScheduleDag;
  int nlevels = 0;
  SolveDev d;
  std::vector<int32_t> h_small_ptr, h_fx_ptr, h_fu_ptr, h_br_ptr, h_bx_ptr, h_fm_ptr, h_bm_ptr;
};

// x (length n, permuted order) is overwritten by (L D L^T)^{-1} x.
// uvec: scratch of sum(nrow) entries of TX.
template <class TF, class TD, class TX>
void ldl_solve(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st);
// ldl_solve with the work areas, counters and error flag of `d` in place of
// dp.d's (the same task lists and offsets)
template <class TF, class TD, class TX>
void ldl_solve_on(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, const SolveDev& d, TX* x, TX* uvec,
                  hipStream_t st);
// nrhs right-hand sides x + r dp.d.sx at once (each tile of the factor read
// once for up to 8 of them); uvec and the partial-sum areas of dp.d hold 8
// right-hand sides at strides dp.d.su, .syp, .sup, .srp, .sxp. With gate
// set, every kernel returns at once if *gate == 0 (read at kernel start).
template <class TF, class TD, class TX>
void ldl_solve_multi(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, int nrhs, TX* uvec,
                     hipStream_t st, const int* gate = nullptr);

// ---------------------------------------------------------------- FGPILU
// Device view of the ILU structure (host/ilu.hpp): U by columns, L by rows
// on the same index arrays, U by rows, level sets of both solves.
struct IluDev {
  int n = 0;
  int64_t nnz = 0;
  const int32_t* Up = nullptr;
  const int32_t* Ui = nullptr;
  const int32_t* col = nullptr;
  const int32_t* src = nullptr;
  const int32_t* ur_ptr = nullptr;
  const int32_t* ur_col = nullptr;
  const int32_t* ur_idx = nullptr;
  const int32_t* lo_rows = nullptr;
  const int32_t* up_rows = nullptr;
  const int* gate = nullptr;  // when set: the triangular solves run only if *gate != 0
};
struct IluLevels {
  std::vector<int32_t> h_lo_ptr, h_up_ptr;
};
// av = A on the pattern (from the scaled values ax), U = L = av, Di = 1 / diag
template <class T>
void ilu_init(const IluDev& g, const double* ax, T* av, T* Ux, T* Lx, T* Di, hipStream_t st);
// one sweep in the reference's order, run level by level over the columns:
// the exact ILU factors of the pattern, in place
template <class T>
void ilu_factor(const IluDev& g, const IluLevels& lv, const T* av, T* Ux, T* Lx, T* Di, hipStream_t st);
// *out = the nonlinear residual; part: one double per 256 entries
template <class T>
void ilu_nlres(const IluDev& g, const T* av, const T* Lx, const T* Ux, double* part, double* out, hipStream_t st);
// x := (L U)^{-1} x (FP64 arithmetic)
template <class T>
void ilu_solve(const IluDev& g, const IluLevels& lv, const T* Lx, const T* Ux, const T* Di, double* x, hipStream_t st);

// ---------------------------------------------------------------- vectors / SpMV
struct CsrDev {
  int32_t n = 0;
  const int32_t* p = nullptr;
  const int32_t* j = nullptr;
  const double* x = nullptr;
  const int* gate = nullptr;  // when set: the product runs only if *gate != 0
};

// y = A x (FP64, rows summed left to right, no fused multiply-add)
void spmv_f64(const CsrDev& A, const double* x, double* y, hipStream_t st);
// r = b - A x
void residual_f64(const CsrDev& A, const double* x, const double* b, double* r, hipStream_t st);

}  // namespace mpir
