// GMRES-IR solver with an LDL^T preconditioner, resident on one gfx950 GPU.
// Host-side mirror of GmresLDLIR<UF,UW,UR> (include/MPIR/gmres_ir.hpp:90-315
// of yex33/MPIR) in its LDL^T form (docs/Design/SoftDetailedDes/MIS.tex:505-568).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "batch_plan.hpp"
#include "common.hpp"
#include "kernels/kernels.hpp"
#include "plan.hpp"

namespace mpir {

struct GmDev;  // device state of one PrecondGmres call (gmres.hip)

// Outcome of a batch solve: per system s and right-hand side r, entry
// s * max_rhs + r of the per-instance arrays (slots r >= nrhs[s] unused).
struct BatchStats {
  int32_t nsys = 0, max_rhs = 0;
  std::vector<int32_t> positive_pivots, zero_pivots;  // per system
  std::vector<int32_t> ir_iterations, gmres_iterations, stop_reason;  // per instance
  std::vector<double> final_rel_residual, done_ms;    // per instance; done_ms from the start of the solve
  double factor_ms = 0.0, solve_ms = 0.0;
  int32_t arnoldi_steps = 0, precond_sweeps = 0;      // lock-step counts of the whole batch
  int32_t graph_nodes = 0;  // launches of the replayed HIP graph (0: the solve was launched kernel by kernel)
  // analyze_batch: joining the systems' plans (host) / device copies, allocation and the wait for them
  double join_ms = 0.0, upload_ms = 0.0;
};

// Device vectors of one batched solve (R right-hand-side slots of N rows;
// io_in / io_out: the caller's right-hand sides and solutions, packed) and
// the GMRES cycle length M.
struct BatchRun {
  int R = 0, M = 0;
  double *vb = nullptr, *vx = nullptr, *vr = nullptr, *vd = nullptr, *vt = nullptr;
  double *io_in = nullptr, *io_out = nullptr;
};

enum Ordering : int { kOrderNatural = 0, kOrderAmd = 1, kOrderNd = 2 };

// Dnrm2 of the two-vector {a, b}, as PrecondGmres forms the Givens norm
// (include/MPIR/gmres_ir.hpp:702; ops.hpp:231-255: scaled sum of squares,
// last entry first)
__host__ __device__ inline double dnrm2_pair(double a, double b) {
  double scale = 0.0, ssq = 1.0;
  const double v[2] = {b, a};
  for (double x : v) {
    if (x != 0.0) {
      const double ax = x < 0 ? -x : x;
      if (scale < ax) {
        ssq = 1.0 + ssq * ((scale / ax) * (scale / ax));
        scale = ax;
      } else {
        ssq = ssq + ((ax / scale) * (ax / scale));
      }
    }
  }
  return scale * __builtin_sqrt(ssq);
}

// preconditioner M of GMRES-IR
constexpr int kPrecLdl = 0;  // L D L^T of the reordered matrix (the default)
constexpr int kPrecIlu = 1;  // FGPILU of the matrix in its given order, the reference's shipped Compute

// scaling of K before the factorization
constexpr int kScaleNone = 0;
constexpr int kScaleDiag = 1;  // the reference's s_i = 1/sqrt|k_ii| (gmres_ir.hpp:347-363)
constexpr int kScaleRuiz = 2;  // symmetric Ruiz equilibration of the max-norms
constexpr int kRuizSweeps = 10;
constexpr double kRuizTol = 1e-2;

struct SolveStats {
  int32_t n = 0;
  int64_t nnz = 0;
  int64_t nnz_l = 0;
  int32_t nsupernodes = 0;
  int32_t nlevels = 0;
  int32_t max_front = 0;
  double factor_flops = 0.0;
  int32_t positive_pivots = -1;
  int32_t ir_iterations = 0;
  int32_t gmres_iterations = 0;  // summed over all PrecondGmres calls
  int32_t precond_applications = 0;
  int32_t stop_reason = 0;        // 0 max iterations, 1 small update, 2 stagnation, 3 zero residual
  double final_rel_residual = 0;  // ||r||_inf / ||b_s||_inf in the scaled system
  double analyze_ms = 0, factor_ms = 0, solve_ms = 0;
  int64_t device_bytes = 0;
  int32_t ruiz_sweeps = 0;  // Ruiz sweeps of the last factorization
  int32_t ilu_sweeps = 0;   // FGPILU: sweeps of the last factorization
  double ilu_nl_residual0 = 0.0, ilu_nl_residual = 0.0;  // FGPILU: nonlinear residual before / after
  int32_t preconditioner = kPrecLdl;
};

// Record of the last GMRES-IR solve, in the terms of the reference's
// VERBOSE log (gmres_ir.hpp:539-631)
struct SolveLog {
  struct Iter {
    int32_t iteration;
    double residual, d_norm, timestamp_ms, inner_time_ms;
  };
  double initial_residual = 0.0;
  std::vector<Iter> iterations;
  int32_t stop_reason = 0, total_iterations = 0;
  double total_time_ms = 0.0;
};

// Outcome of LUIR_QDLDL (include/qdldl-luir/luir.hpp:33-241, stats.hpp)
struct LuirResult {
  int32_t steps = 0;          // num_refinement_steps_
  double rel_residual = 0.0;  // relative_residual_
  double time_solve = 0.0, time_ref = 0.0;  // ms
};

// What GmresLdl::autotune chose and measured (ms: HIP events on the
// handle's stream, averages over the requested runs).
struct AutotuneResult {
  FactorTuning factor;
  SolveTuning solve;
  int schedule = 0;  // kScheduleLevels or kScheduleDag
  double factor_ms = 0.0, precond_ms = 0.0;                  // chosen
  double factor_ms_before = 0.0, precond_ms_before = 0.0;    // the tuning in place before the call
  int32_t candidates = 0;
};

class GmresLdl {
 public:
  GmresLdl(int uf, int uw, int ur);
  ~GmresLdl();
  GmresLdl(const GmresLdl&) = delete;
  GmresLdl& operator=(const GmresLdl&) = delete;

  void set_max_ir(int64_t v) { max_ir_ = v; }
  void set_max_gmres(int64_t v) { max_gmres_ = v; }
  void set_tol(double v) { tol_ = v; }
  void set_ordering(int o) { ordering_ = o; }
  // diagonal scaling of Compute (gmres_ir.hpp:347-363); off for the bare
  // QDLDL_factor drop-in
  // 0 none, 1 the reference's diagonal scaling (default), 2 Ruiz
  void set_scaling(int mode) {
    MPIR_CHECK(mode >= kScaleNone && mode <= kScaleRuiz, "scaling mode must be 0, 1 or 2");
    scaling_ = mode;
  }
  // supernode amalgamation: 0 none, 1 .. 3 (default; relax_supernodes);
  // takes effect at the next analyze
  // panel-step launches of at least min_tasks tasks are split (0: never);
  // the factor does not change
  void set_panel_split(int min_tasks) {
    MPIR_CHECK(min_tasks >= 0, "panel split threshold must be >= 0");
    pan_split_min_ = min_tasks;
    dfp_.pan_split_min = min_tasks;
  }
  void set_relax(int level) {
    MPIR_CHECK(level >= 0 && level <= 3, "amalgamation level must be 0 .. 3");
    relax_ = level;
  }
  int ordering() const { return ordering_; }
  int relax() const { return relax_; }
  // schedule of the triangular solves: kScheduleLevels or kScheduleDag
  void set_schedule(int sched) {
    MPIR_CHECK(sched == kScheduleLevels || sched == kScheduleDag, "solve schedule must be 0 (levels) or 1 (dag)");
    schedule_ = sched;
    dsp_.schedule = sched;
  }
  // schedule of the numeric factorization: kScheduleLevels (default; one
  // launch per level and step) or kScheduleDag (the levels above the small
  // fronts in one dependency-driven launch, with an update-matrix pool that
  // reuses no block there); both give the same factor bit for bit. Selecting
  // the DAG on an analyzed handle lays the pool out again (factor again).
  void set_factor_schedule(int sched);
  int factor_schedule() const { return factor_schedule_; }
  // the dependency-driven launch of the current plan (state error before analyze)
  void factor_schedule_info(int32_t* dag_level, int32_t* nlevels, int64_t* tasks) const {
    if (!plan_) throw Error(kErrState, "analyze first");
    *dag_level = plan_->fdag_level;
    *nlevels = plan_->sn.nlevels;
    *tasks = static_cast<int64_t>(plan_->fdag.size());
  }
  // kPrecLdl (default) or kPrecIlu with the ILU level and the sweep limit of
  // the reference's Compute (its ILU_k argument and MAX_SWEEPS parameter);
  // takes effect at the next analyze
  void set_preconditioner(int kind, int32_t ilu_k, int32_t max_sweeps);
  int preconditioner() const { return precond_; }
  // FGPILU factors in the layout of the reference (U by columns, L by rows on
  // the same indices), converted to double: sizes n+1, nnz, nnz, nnz, n
  void export_ilu(int32_t* Up, int32_t* Ui, double* Ux, double* Lx, double* UDinv);
  int64_t ilu_nnz() const { return static_cast<int64_t>(ilu_Ui_.size()); }
  // Waits for the handle's stream and reports a triangular solve whose
  // dependency wait expired.
  void sync_check();

  // Symbolic phase on the pattern of the upper CSC (reusable across values).
  void analyze(int32_t n, const int32_t* Ap, const int32_t* Ai);
  // Numeric phase: scaling, conversion, factorization. Ax in the original
  // entry order of the analyzed pattern; host or device pointer.
  void factor(const double* Ax, bool device_ptr);
  // x = K^{-1} b for nrhs right-hand sides with leading dimension ld.
  void solve(const double* b, double* x, int nrhs, int64_t ld, bool device_ptr);

  const SolveStats& stats() const { return stats_; }
  std::string log_json() const;  // the log of the last solve (JSON text)
  hipStream_t stream() const { return stream_; }
  bool analyzed() const { return analyzed_; }
  bool factored() const { return factored_; }
  int32_t n() const { return n_; }
  const std::vector<int32_t>& perm() const { return perm_; }

  // ---- batches of independent systems, solved as one block-diagonal system
  // Symbolic phase of every system from the pattern cache (analysed on the
  // fly when missing), combined into one plan over the forest of their
  // elimination trees.
  void analyze_batch(const std::vector<std::shared_ptr<const PatternEntry>>& parts);
  // values of every system (original entry order of its pattern)
  void factor_batch(const double* const* Ax, bool device_ptr);
  // nrhs[s] right-hand sides for system s, column-major with ld = n_s
  void solve_batch(const int32_t* nrhs, const double* const* b, double* const* x, bool device_ptr);
  const BatchStats& batch_stats() const { return bstats_; }
  bool is_batch() const { return static_cast<bool>(batch_); }
  // whether repeated batch solves replay a captured HIP graph (default on)
  void set_batch_graph(bool on) { batch_graph_ = on; }
  int32_t batch_size() const;
  int32_t batch_n(int32_t s) const;

  // classic iterative refinement of LUIR_QDLDL on the current factor
  // (natural ordering, no scaling); b, x host arrays
  void luir(const double* b, double* x, int maxiter, LuirResult* out);
  // building blocks exposed for tests and the kernel-level C entry points
  void precond_apply_dev(double* x);  // x := (L D L^T)^{-1} x, permuted order
  // Bytes of work memory one application of M^{-1} needs (counters, error
  // flag, partial sums, update vectors), and the application with that work
  // memory carved from `work` (device memory, 256-byte aligned) instead of
  // the handle's buffers; waits for the result (noalloc.hip).
  int64_t precond_work_bytes() const;
  void precond_apply_in(double* x, void* work);
  // the same for R vectors of n entries, contiguous (the factor read once
  // for up to 8 of them)
  void precond_apply_slots(double* x, int R);
  void spmv_dev(const double* x, double* y);  // y = K_s x, permuted, scaled
  // factor (permuted order) copied out in QDLDL's CSC layout
  void export_factor(int64_t* Lp, int32_t* Li, void* Lx, void* D, void* Dinv, int prec);
  // Average time (ms) of `reps` back-to-back runs of one kernel group on the
  // handle's stream, measured with HIP events: 0 SpMV, 1 M^{-1} application,
  // 2 numeric factorization (values already loaded by factor()).
  double time_kernels(int what, int reps);
  // one factorization of the current values with the dependency-driven
  // launch's task timeline (4 entries per task) in out[0 .. cap); returns
  // the number of tasks (cap = 0: only that)
  int64_t trace_factor(uint64_t* out, int64_t cap);
  // one M^{-1} application (dependency-driven schedule) with the task
  // timeline of both passes (4 entries per task, forward then backward) in
  // out[0 .. cap); returns the number of tasks, *nfwd those of the forward pass
  int64_t trace_precond(uint64_t* out, int64_t cap, int64_t* nfwd);
  // per supernode of the analyzed plan (the indices the traces record): its
  // level in the supernodal tree, columns and rows below its diagonal block,
  // written when cap >= the number of supernodes; returns that number
  int32_t supernode_info(int32_t* level, int32_t* ncols, int32_t* nrows, int32_t cap) const;
  // Times the factor's update grouping and super-block edge and the
  // triangular solves' narrow-front limit and schedule on this handle's
  // factor (reps runs per candidate), keeps the fastest of each and leaves
  // the factor recomputed with it. The choice stays with the handle: later
  // analyses (analyze, analyze_batch) build their plans with it.
  AutotuneResult autotune(int reps);
  // Sets the tunings that autotune would choose; the plan in place is
  // re-planned with them (an existing factor stays valid: the storage layout
  // does not change).
  void set_tuning(const FactorTuning& f, const SolveTuning& s);
  // average ms of one M^{-1} application to `rhs` vectors at once
  double time_precond_slots(int rhs, int reps);
  // Load a factor given in QDLDL's CSC layout (host arrays, Lx/Dinv in the
  // handle's factor precision) for use as M^{-1} only.
  void load_factor_csc(int32_t n, const int64_t* Lp, const int32_t* Li, const void* Lx, const void* Dinv);
  int64_t plan_builds() const { return plan_builds_; }

 private:
  void solve_one(const double* b_dev, double* x_dev);
  void factor_loaded(double t0);  // numeric phase on the values in d_ax_
  void alloc_vectors();  // the FP64 vectors, A's values and the flags every preconditioner uses
  void analyze_ilu(int32_t n, const int32_t* Ap, const int32_t* Ai, double t0);
  void factor_ilu();
  template <class T>
  void factor_ilu_t();
  double ilu_nlres_now(const void* L, const void* U);
  void batch_count_pivots();
  void batch_collect_pivots(bool zero_pivot);
  // the launch sequence of one batched GMRES-IR solve, every scalar on the
  // device (batch.hip); the same sequence runs eagerly or is captured
  void batch_enqueue(const BatchRun& br);
  void set_solve_slots(int slots);
  double* bpinned(size_t count);
  // Host arrays to device buffers. While staging_ is set (a batch's
  // uploads) the copies are recorded and flush_staged() runs them through
  // one pinned buffer: host copies into it on several threads, then
  // asynchronous DMA (the runtime stages copies from pageable memory one at
  // a time, which serialised the serving lanes). The arrays must not change
  // before the flush; the caller waits for the stream before the pinned
  // buffer is reused.
  template <class T>
  // While staging_ is set (a batch's analysis), the copy is made when the
  // stage is flushed, from v's storage then: v must outlive flush_staged().
  void up(DevBuf& b, const std::vector<T>& v) {
    b.alloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (v.empty()) return;
    if (staging_) staged_.push_back({b.get(), v.data(), v.size() * sizeof(T)});
    else MPIR_HIP(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
  }
  void flush_staged();
  // device-resident PrecondGmres (device vectors, permuted order), skipped
  // when *gate == 0 (gate may be null)
  void precond_gmres(const double* b, double* res, const int* gate);
  GmDev gm_dev(int m);
  void norm_to(const double* v, int op, double* out, const int* gate);
  void precond_apply_gated(double* x, const int* gate);
  void wait_mirror(int slot, double target);
  double dnrm2(const double* v);
  double infnrm(const double* v);
  void release_numeric();
  void upload_plan();
  void upload_factor_dag();
  // rebuilds plan_'s step and solve schedules for ftune_ / stune_ where they
  // differ from the plan's
  void apply_tuning();
  template <class T>
  FactorArgs<T> factor_args() {
    return FactorArgs<T>{d_first_.as<int32_t>(), d_nrow_.as<int32_t>(), d_rows_off_.as<int64_t>(),
                         d_rows_.as<int32_t>(), d_relidx_.as<int32_t>(), d_panel_off_.as<int64_t>(),
                         d_upd_off_.as<int64_t>(), d_child_ptr_.as<int32_t>(), d_child_list_.as<int32_t>(),
                         d_panel_.as<T>(), d_upool_.as<T>(), d_D_.as<T>(), d_Dinv_.as<T>(), d_err_.as<int>(),
                         dfp_.zbuf_fronts > 0 ? d_zbuf_.as<T>() : nullptr, d_zslot_.as<int32_t>(),
                         dfp_.zbuf_fronts};
  }
  template <class TF, class TD>
  SolveArgs<TF, TD> solve_args(void* panel) {
    return SolveArgs<TF, TD>{d_first_.as<int32_t>(), d_nrow_.as<int32_t>(), d_rows_off_.as<int64_t>(),
                             d_rows_.as<int32_t>(), d_relidx_.as<int32_t>(), d_panel_off_.as<int64_t>(),
                             d_child_ptr_.as<int32_t>(), d_child_list_.as<int32_t>(), static_cast<const TF*>(panel),
                             d_Dinv_.as<TD>()};
  }
  // storage rounding of the factor's working values (see factor_numeric)
  int rounding() const { return uf_ == kF16 ? 1 : uf_ == kBF16 ? 2 : 0; }
  void round_factor16();  // the 16-bit copy of the factor (and its range check)
  template <class TF>
  void luir_impl(const double* b, double* x, int maxiter, LuirResult* out);
  void alloc_numeric();

  int uf_, uw_, ur_;
  int64_t max_ir_ = 10, max_gmres_ = 10;
  double tol_ = 1e-10;
  int ordering_ = kOrderAmd;
  int scaling_ = kScaleDiag;
  int schedule_ = kScheduleDag;
  int factor_schedule_ = kScheduleLevels;
  FactorTuning ftune_;
  SolveTuning stune_;
  int relax_ = kDefaultAmalgamation;  // relaxed supernodes, level 0 .. 3 (AMD / ND orderings only)
  int pan_split_min_ = kPanelSplitMin;
  bool relaxed_ = false;  // the current plan uses them
  std::vector<int32_t> Cp_, Ci_, parent_, Lnz_;  // pattern of P K P^T, etree
  int64_t true_nnz_l_ = 0;
  double true_flops_ = 0.0;
  bool analyzed_ = false, factored_ = false, precond_only_ = false;
  // pattern of the factor given to load_factor_csc, and how many plans the
  // handle has built from such patterns
  std::vector<int64_t> loaded_Lp_;
  std::vector<int32_t> loaded_Li_;
  int64_t plan_builds_ = 0;
  hipStream_t stream_ = nullptr;
  int32_t n_ = 0;
  int64_t nnz_ = 0;
  SolveStats stats_;
  SolveLog log_;

  // pattern
  std::vector<int32_t> Ap_, Ai_, perm_;
  std::unique_ptr<FactorPlan> plan_;
  DeviceFactorPlan dfp_;
  DeviceSolvePlan dsp_;
  std::vector<DevBuf> keep_;  // device copies of plan arrays
  DevBuf d_first_, d_nrow_, d_rows_off_, d_rows_, d_relidx_, d_panel_off_, d_upd_off_, d_child_ptr_, d_child_list_;
  DevBuf d_zbuf_, d_zslot_;  // split panel steps: Z_kk slots, a front's slot
  // host copy of the slot map: a staged upload reads it when the stage is
  // flushed, so it must outlive upload_plan
  std::vector<int32_t> h_zslot_;
  // split panel steps: each step's owner tasks and other tasks (uploaded
  // from these members, which outlive a staged flush)
  std::vector<Task4> h_pan_own_, h_pan_rest_;
  DevBuf d_pan_own_, d_pan_rest_;
  DevBuf d_csr_p_, d_csr_j_, d_csr_src_, d_csr_x_;
  DevBuf d_entry_col_, d_entry_row_, d_diag_pos_, d_perm_;
  // numeric
  DevBuf d_ax_, d_axs_, d_ad_, d_adp_, d_panel_, d_panel16_, d_upool_, d_D_, d_Dinv_, d_err_, d_cnt_, d_nonfinite_;
  DevBuf d_uvec_, d_scratch_, d_serr_, d_counters_;
  DevBuf d_fdag_, d_fdag_meta_, d_fcounters_, d_ftrace_, d_strace_;  // dependency-driven factorization
  // GMRES workspace
  DevBuf d_V_, d_w_, d_b_, d_x_, d_r_, d_d_, d_red_, d_small_;
  DevBuf d_hcol_, d_givens_;
  DevBuf d_gm_, d_gmi_, d_ir_;  // device state of GMRES-IR (scalars, Hessenberg, rotations, gates)
  int64_t step_seq_ = 0, ir_seq_ = 0;  // ordinals of the records the device publishes
  std::vector<double> h_small_;
  double* h_pinned_ = nullptr;
  double* h_mirror_ = nullptr;  // h_pinned_ as the device addresses it
  void wait_stream();
  // FGPILU
  int precond_ = kPrecLdl;
  int32_t ilu_k_ = 0, ilu_max_sweeps_ = 10;
  IluDev ilu_g_;
  IluLevels ilu_lv_;
  std::vector<int32_t> ilu_Up_, ilu_Ui_;
  DevBuf d_ilu_idx_[9];  // Up, Ui, col, src, ur_ptr, ur_col, ur_idx, lo_rows, up_rows
  DevBuf d_ilu_av_, d_ilu_u_, d_ilu_l_, d_ilu_d_, d_ilu_part_, d_ilu_out_;
  // batch mode
  std::unique_ptr<BatchSystem> batch_;
  std::vector<std::shared_ptr<const PatternEntry>> batch_parts_;
  BatchStats bstats_;
  DevBuf d_b_sys_of_row_, d_b_col_base_, d_b_lperm_, d_b_ch_start_, d_b_ch_len_, d_b_sys_ch_;
  DevBuf d_b_off_, d_b_nrhs_, d_b_piv_;
  DevBuf d_b_vec_;  // b, x, r, d, w of every slot
  DevBuf d_bs_;     // per-instance state of the device-resident batched GMRES-IR
  int32_t b_nch_ = 0, b_slots_ = 0;
  // the captured solve: replayed whenever a solve's key (shapes,
  // tolerances, plan, buffers) matches it; captured when a solve repeats the
  // key of the solve before it (a one-off batch runs eagerly)
  hipGraphExec_t bgraph_ = nullptr;
  int32_t bgraph_nodes_ = 0;
  std::vector<uint64_t> bgraph_key_, blast_key_;
  bool batch_graph_ = true;
  uint64_t plan_epoch_ = 0;  // counts upload_plan calls
  int solve_slots_ = 0;  // right-hand sides the solve work areas hold
  double* b_pinned_ = nullptr;
  size_t b_pinned_count_ = 0;
  struct Staged {
    void* dst;
    const void* src;
    size_t bytes;
  };
  bool staging_ = false;
  std::vector<Staged> staged_;
  char* stage_pinned_ = nullptr;
  size_t stage_cap_ = 0;
};

// x := M^{-1} x on the current device with the work memory of the solves
// in a static device array: no device allocation (noalloc.hip)
void precond_apply_noalloc(GmresLdl& s, double* x);

}  // namespace mpir
