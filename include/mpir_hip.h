/*
 * mpir_hip.h -- C ABI of the MI355X (gfx950) implementation of the
 * yex33/MPIR solve path for quasi-definite systems K = [H A^T; A -D]:
 * fill-reducing ordering, QDLDL elimination tree, LDL^T factorization and
 * GMRES-IR with LDL^T as the preconditioner.
 *
 * Plain C types only. Index arrays are int32 ("_i32") or int64 ("_i64");
 * int64 entry points narrow to int32 internally and fail with
 * MPIR_ERR_INVALID when a value does not fit.
 *
 * Every function is synchronous with respect to the host unless its name
 * ends in "_dev" (device pointers, work ordered on the solver's stream and
 * complete on return). Errors: negative status; mpir_last_error() returns a
 * message for the calling thread.
 */
#ifndef MPIR_HIP_H
#define MPIR_HIP_H

#include <stddef.h>
#include <stdint.h>

/* The library is built with hidden symbol visibility; only the functions
 * declared here are exported. */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* precision tags */
#define MPIR_F64 0
#define MPIR_F32 1
#define MPIR_F16 2
#define MPIR_BF16 3

/* status codes (QDLDL's own -1 / -2 keep their meaning where they apply) */
#define MPIR_OK 0
#define MPIR_ERR_INVALID (-100)
#define MPIR_ERR_HIP (-101)
#define MPIR_ERR_NODEVICE (-102)
#define MPIR_ERR_STATE (-103)
#define MPIR_ERR_NUMERIC (-104)
#define MPIR_ERR_INTERNAL (-105)

/* scalings */
#define MPIR_SCALE_NONE 0
#define MPIR_SCALE_DIAG 1
#define MPIR_SCALE_RUIZ 2

/* orderings */
#define MPIR_ORDER_NATURAL 0 /* the matrix as given (QDLDL's behaviour) */
#define MPIR_ORDER_AMD 1     /* approximate minimum degree + etree postorder */
#define MPIR_ORDER_ND 2      /* nested dissection (AMD on small parts) + postorder */

const char* mpir_last_error(void);
const char* mpir_version(void);
/* 1 if a gfx950 device is usable by this process, 0 otherwise */
int mpir_device_ok(void);

/* ------------------------------------------------------------------ QDLDL
 * Replaces QDLDL_etree<INT> (yex33/MPIR include/MPIR/qdldl.hpp:48-105).
 * Same arguments, same outputs, same return: sum(Lnz), -1 for an empty
 * column or an entry below the diagonal, -2 when sum(Lnz) overflows INT.
 * Host computation (symbolic, pointer chasing). */
int32_t mpir_qdldl_etree_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* work, int32_t* Lnz,
                             int32_t* etree);
int64_t mpir_qdldl_etree_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, int64_t* work, int64_t* Lnz,
                             int64_t* etree);

/* Replaces QDLDL_factor<INT,PREC> (include/MPIR/qdldl.hpp:172-345). Same
 * arguments and outputs (Lp, Li, Lx in CSC with rows ascending, D, Dinv);
 * returns the number of positive entries of D, or -1 on a zero pivot.
 * Deviation: a factor with a pivot that is not finite (an overflow or a NaN
 * in Ax) also returns -1, where the reference returns it with Inf/NaN
 * entries. The factorization runs on the GPU; bwork/iwork/fwork are accepted
 * for signature compatibility and may be NULL. Like the reference's, these
 * functions are reentrant: each calling thread has its own device state, so
 * threads may call them concurrently. */
int32_t mpir_qdldl_factor_f64_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, int32_t* Lp,
                                  int32_t* Li, double* Lx, double* D, double* Dinv, const int32_t* Lnz,
                                  const int32_t* etree, unsigned char* bwork, int32_t* iwork, double* fwork);
int32_t mpir_qdldl_factor_f32_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const float* Ax, int32_t* Lp,
                                  int32_t* Li, float* Lx, float* D, float* Dinv, const int32_t* Lnz,
                                  const int32_t* etree, unsigned char* bwork, int32_t* iwork, float* fwork);
int64_t mpir_qdldl_factor_f64_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax, int64_t* Lp,
                                  int64_t* Li, double* Lx, double* D, double* Dinv, const int64_t* Lnz,
                                  const int64_t* etree, unsigned char* bwork, int64_t* iwork, double* fwork);
int64_t mpir_qdldl_factor_f32_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const float* Ax, int64_t* Lp,
                                  int64_t* Li, float* Lx, float* D, float* Dinv, const int64_t* Lnz,
                                  const int64_t* etree, unsigned char* bwork, int64_t* iwork, float* fwork);

/* QDLDL_factor<INT,PREC> with a 16-bit PREC: Ax, Lx, D and Dinv hold IEEE
 * binary16 (f16) or bfloat16 (bf16) bit patterns. The factor is computed
 * with FP32 arithmetic, every matrix value and every entry of L rounded to
 * the format as it is stored (round to nearest even); D and Dinv are
 * rounded on output. Returns the number of positive entries of D, or -1 on
 * a zero pivot or a value outside the format's range. */
int32_t mpir_qdldl_factor_f16_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const uint16_t* Ax, int32_t* Lp,
                                  int32_t* Li, uint16_t* Lx, uint16_t* D, uint16_t* Dinv, const int32_t* Lnz,
                                  const int32_t* etree, unsigned char* bwork, int32_t* iwork, uint16_t* fwork);
int32_t mpir_qdldl_factor_bf16_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const uint16_t* Ax, int32_t* Lp,
                                   int32_t* Li, uint16_t* Lx, uint16_t* D, uint16_t* Dinv, const int32_t* Lnz,
                                   const int32_t* etree, unsigned char* bwork, int32_t* iwork, uint16_t* fwork);
int64_t mpir_qdldl_factor_f16_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const uint16_t* Ax, int64_t* Lp,
                                  int64_t* Li, uint16_t* Lx, uint16_t* D, uint16_t* Dinv, const int64_t* Lnz,
                                  const int64_t* etree, unsigned char* bwork, int64_t* iwork, uint16_t* fwork);
int64_t mpir_qdldl_factor_bf16_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const uint16_t* Ax, int64_t* Lp,
                                   int64_t* Li, uint16_t* Lx, uint16_t* D, uint16_t* Dinv, const int64_t* Lnz,
                                   const int64_t* etree, unsigned char* bwork, int64_t* iwork, uint16_t* fwork);

/* Replaces QDLDL_solve<INT,FACTPREC,XPREC> (include/MPIR/qdldl.hpp:417-432):
 * x := (L D L^T)^{-1} x with L in FACTPREC and arithmetic in XPREC. The
 * solve runs on the GPU. Returns 0, or a negative status (the reference's
 * solve is void and unchecked). */
int32_t mpir_qdldl_solve_f64_f64_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const double* Lx,
                                     const double* Dinv, double* x);
int32_t mpir_qdldl_solve_f32_f64_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const float* Lx,
                                     const float* Dinv, double* x);
int32_t mpir_qdldl_solve_f32_f32_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const float* Lx,
                                     const float* Dinv, float* x);
int64_t mpir_qdldl_solve_f64_f64_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const double* Lx,
                                     const double* Dinv, double* x);
int64_t mpir_qdldl_solve_f32_f64_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const float* Lx,
                                     const float* Dinv, double* x);
int64_t mpir_qdldl_solve_f32_f32_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const float* Lx,
                                     const float* Dinv, float* x);
/* QDLDL_solve with a 16-bit FACTPREC (bit patterns as above), XPREC float or
 * double: the factor's values are exact in FP32 and the solve computes in
 * XPREC arithmetic as the FP32 forms do */
int32_t mpir_qdldl_solve_f16_f32_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx,
                                     const uint16_t* Dinv, float* x);
int32_t mpir_qdldl_solve_f16_f64_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx,
                                     const uint16_t* Dinv, double* x);
int64_t mpir_qdldl_solve_f16_f32_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx,
                                     const uint16_t* Dinv, float* x);
int64_t mpir_qdldl_solve_f16_f64_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx,
                                     const uint16_t* Dinv, double* x);
int32_t mpir_qdldl_solve_bf16_f32_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx,
                                      const uint16_t* Dinv, float* x);
int32_t mpir_qdldl_solve_bf16_f64_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx,
                                      const uint16_t* Dinv, double* x);
int64_t mpir_qdldl_solve_bf16_f32_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx,
                                      const uint16_t* Dinv, float* x);
int64_t mpir_qdldl_solve_bf16_f64_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx,
                                      const uint16_t* Dinv, double* x);

/* ------------------------------------------------------------------ LUIR
 * Replaces LUIR_QDLDL<FACTPREC, double, double, double, INT> and its Stats
 * (include/qdldl-luir/luir.hpp:33-241, stats.hpp:6-114): QDLDL etree and
 * factor of the matrix as given (upper CSC, values rounded to FACTPREC),
 * x0 = a FACTPREC solve, then classic iterative refinement with the residual
 * in double, each correction solved in FACTPREC, stopping when
 * ||d||/||x|| <= 10 eps(double) or ||d||/||d_prev|| >= 0.5, at most maxiter
 * steps (maxiter = 0: one solve in double). Factor and refinement run on the
 * GPU. Returns 0, -1 / -2 where QDLDL_etree or QDLDL_factor would, or a
 * negative status; the reference exits the process instead. */
typedef struct mpir_luir_stats {
  int32_t n;
  int64_t nnz;                  /* entries of triu(A) */
  int64_t lnz;                  /* entries of L below the diagonal */
  int32_t num_refinement_steps;
  double relative_residual;     /* ||b - A x||_inf / ||b||_inf, residual rounded to FACTPREC */
  double time_sym, time_num_fact, time_solve, time_ref;  /* seconds */
} mpir_luir_stats;
int32_t mpir_luir_qdldl_f64_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b,
                                double* x, mpir_luir_stats* stats, int32_t maxiter);
int32_t mpir_luir_qdldl_f32_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b,
                                double* x, mpir_luir_stats* stats, int32_t maxiter);
int64_t mpir_luir_qdldl_f64_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax, const double* b,
                                double* x, mpir_luir_stats* stats, int32_t maxiter);
int64_t mpir_luir_qdldl_f32_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax, const double* b,
                                double* x, mpir_luir_stats* stats, int32_t maxiter);

/* ------------------------------------------------------------------ GMRES-IR
 * Replaces GmresLDLIR<UF,UW,UR> (include/MPIR/gmres_ir.hpp:90-315) in its
 * LDL^T form (docs/Design/SoftDetailedDes/MIS.tex:505-568). */
typedef struct mpir_gmres_ldl* mpir_gmres_ldl_t;

typedef struct mpir_gmres_stats {
  int32_t n;
  int64_t nnz;                 /* upper-triangular input entries */
  int64_t nnz_l;               /* strictly-lower entries of L */
  int32_t nsupernodes;
  int32_t nlevels;             /* levels of the supernodal tree */
  int32_t max_front;
  double factor_flops;
  int32_t positive_pivots;     /* QDLDL_factor's return value */
  int32_t ir_iterations;       /* last solve */
  int32_t gmres_iterations;    /* last solve, summed over PrecondGmres calls */
  int32_t precond_applications;
  int32_t stop_reason;         /* 0 max iterations, 1 small update, 2 stagnation, 3 zero residual */
  double final_rel_residual;   /* ||b_s - K_s x||_inf / ||b_s||_inf, scaled system */
  double analyze_ms, factor_ms, solve_ms;
  int64_t device_bytes;
  int32_t ruiz_sweeps;         /* equilibration sweeps of the last factorization */
  int32_t preconditioner;      /* MPIR_PREC_LDL or MPIR_PREC_FGPILU */
  int32_t ilu_sweeps;          /* FGPILU: sweeps of the last factorization */
  double ilu_nl_residual0;     /* FGPILU: nonlinear residual of the initial guess */
  double ilu_nl_residual;      /* FGPILU: nonlinear residual after the last sweep */
} mpir_gmres_stats;

/* uf: MPIR_F64, MPIR_F32, or MPIR_F16 / MPIR_BF16 (factor computed with
 * FP32 arithmetic and its working values stored rounded to the 16-bit
 * format; L and the triangular-solve operands held in that format, D^{-1}
 * in FP32); uw, ur: MPIR_F64 */
int mpir_gmres_ldl_create(int uf, int uw, int ur, mpir_gmres_ldl_t* out);
void mpir_gmres_ldl_destroy(mpir_gmres_ldl_t h);
int mpir_gmres_ldl_set_max_ir(mpir_gmres_ldl_t h, int64_t n);        /* SetMaxIRIterations */
int mpir_gmres_ldl_set_max_gmres(mpir_gmres_ldl_t h, int64_t n);     /* SetMaxGmresIterations */
int mpir_gmres_ldl_set_tol(mpir_gmres_ldl_t h, double tol);          /* SetTolerance */
int mpir_gmres_ldl_set_ordering(mpir_gmres_ldl_t h, int ordering);  /* before analyze */
/* scaling of K before factoring: MPIR_SCALE_NONE, MPIR_SCALE_DIAG (the
 * reference's s_i = 1/sqrt|k_ii| where |k_ii| > eps(UF), gmres_ir.hpp:347-363;
 * the default) or MPIR_SCALE_RUIZ (symmetric Ruiz equilibration: up to 10
 * sweeps, until every row's max-norm is within 1e-2 of one) */
int mpir_gmres_ldl_set_scaling(mpir_gmres_ldl_t h, int mode);
/* schedule of the triangular solves inside M^{-1} (no counterpart in the
 * reference, whose QDLDL_solve is a sequential loop): MPIR_SCHED_LEVELS, one
 * launch per task kind and tree level, or MPIR_SCHED_DAG (default), one
 * persistent launch per direction driven by per-front dependency counters.
 * Both give bit-identical results. */
#define MPIR_SCHED_LEVELS 0
#define MPIR_SCHED_DAG 1
int mpir_gmres_ldl_set_schedule(mpir_gmres_ldl_t h, int schedule);
/* supernode amalgamation of the symbolic phase, before analyze (no
 * counterpart in the reference, whose QDLDL_factor works column by column):
 * 0 fundamental supernodes, 1 to 3 (default) relaxed supernodes that merge a
 * child into its parent more freely (fewer, larger fronts and a shallower
 * elimination tree, for explicit zeros stored in the factor). Orderings AMD
 * and ND only. */
int mpir_gmres_ldl_set_amalgamation(mpir_gmres_ldl_t h, int level);
/* panel-step launches of the per-level factorization with at least
 * min_tasks tasks (default 512; 0: never) run as two launches: each
 * front's diagonal block is eliminated once and its explicit inverse
 * published, then the rows below read it back instead of eliminating the
 * block again. The factor is the same bit for bit (no counterpart in the
 * reference, whose QDLDL_factor works column by column). */
int mpir_gmres_ldl_set_panel_split(mpir_gmres_ldl_t h, int min_tasks);
/* schedule of the numeric factorization (no counterpart in the reference,
 * whose QDLDL_factor is a sequential loop over columns, qdldl.hpp:228-342):
 * MPIR_SCHED_LEVELS (default), one launch per level of the supernodal tree
 * and per panel step, or MPIR_SCHED_DAG, the levels above the last small
 * front in one persistent launch whose tile tasks wait on per-front and
 * per-column-block completion counters. Both give bit-identical factors. */
int mpir_gmres_ldl_set_factor_schedule(mpir_gmres_ldl_t h, int schedule);
/* the factorization schedule of an analyzed handle: the schedule in use,
 * the first level of the dependency-driven launch (== nlevels: none), the
 * number of levels and the tasks of that launch (no counterpart in the
 * reference) */
int mpir_gmres_ldl_factor_schedule_info(mpir_gmres_ldl_t h, int* schedule, int32_t* dag_level, int32_t* nlevels,
                                        int64_t* dag_tasks);
/* preconditioner M of GMRES-IR, before analyze: MPIR_PREC_LDL (default),
 * the L D L^T of the reordered matrix, or MPIR_PREC_FGPILU, the reference's
 * shipped GmresLDLIR::Compute (include/MPIR/gmres_ir.hpp:323-531): fine-grained
 * parallel ILU of the diagonally scaled matrix in its given order, on the
 * ILU(ilu_k) pattern (its ILU_k argument, 0..8), sweeping until the nonlinear
 * residual changes by less than the tolerance or max_sweeps (its MAX_SWEEPS);
 * M^{-1} = U^{-1} L^{-1} (TriangularSolve, :743-774). FGPILU factors are f64
 * or f32. The sweeps here are synchronous (every entry from the previous
 * sweep), the reference's run in place, so the two agree once converged. */
#define MPIR_PREC_LDL 0
#define MPIR_PREC_FGPILU 1
int mpir_gmres_ldl_set_preconditioner(mpir_gmres_ldl_t h, int kind, int32_t ilu_k, int32_t max_sweeps);
/* FGPILU factors after factor: U by columns (Up n+1, Ui nnz, Ux nnz), L by
 * rows on the same indices (Lx nnz, unit diagonal stored as 1), 1 / U_jj (n);
 * nnz from mpir_gmres_ldl_ilu_nnz after analyze */
int mpir_gmres_ldl_ilu_nnz(mpir_gmres_ldl_t h, int64_t* nnz);
int mpir_gmres_ldl_export_ilu(mpir_gmres_ldl_t h, int32_t* Up, int32_t* Ui, double* Ux, double* Lx, double* UDinv);
/* symbolic phase on the pattern of the upper CSC (reused by later factors) */
int mpir_gmres_ldl_analyze(mpir_gmres_ldl_t h, int32_t n, const int32_t* Ap, const int32_t* Ai);
/* numeric phase; Ax in the analyzed pattern's entry order (host / device) */
int mpir_gmres_ldl_factor(mpir_gmres_ldl_t h, const double* Ax);
int mpir_gmres_ldl_factor_dev(mpir_gmres_ldl_t h, const double* Ax);
/* Compute(Ap, Ai, Ax): analyze when the pattern differs, then factor */
int mpir_gmres_ldl_compute(mpir_gmres_ldl_t h, int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax);
/* Solve(b) for nrhs columns of length n with leading dimension ld */
int mpir_gmres_ldl_solve(mpir_gmres_ldl_t h, const double* b, double* x, int32_t nrhs, int64_t ld);
int mpir_gmres_ldl_solve_dev(mpir_gmres_ldl_t h, const double* b, double* x, int32_t nrhs, int64_t ld);
int mpir_gmres_ldl_stats(mpir_gmres_ldl_t h, mpir_gmres_stats* out);
/* the fill-reducing permutation chosen by analyze: perm[k] = original index */
int mpir_gmres_ldl_perm(mpir_gmres_ldl_t h, int32_t* perm);
/* the factor in QDLDL's CSC layout, permuted order (sizes n+1 / nnz_l / n) */
int mpir_gmres_ldl_export_factor(mpir_gmres_ldl_t h, int64_t* Lp, int32_t* Li, double* Lx, double* D, double* Dinv);
/* kernel-level entry points on the handle's device vectors (permuted,
 * scaled system), for parity tests and benchmarks */
int mpir_gmres_ldl_precond_dev(mpir_gmres_ldl_t h, double* x);         /* x := M^{-1} x */
int mpir_gmres_ldl_spmv_dev(mpir_gmres_ldl_t h, const double* x, double* y); /* y = K_s x */
/* x := M^{-1} x for nrhs vectors of n entries stored one after the other;
 * an FP32 factor is read once for up to 8 of them. Same results, bit for
 * bit, as nrhs calls of mpir_gmres_ldl_precond_dev. */
int mpir_gmres_ldl_precond_multi_dev(mpir_gmres_ldl_t h, double* x, int32_t nrhs);
/* mpir_gmres_ldl_precond_dev that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The work memory of the triangular solves (completion
 * counters, error flag, partial sums, update vectors) comes from a static
 * device array of 64 MiB that the library holds on each GPU it runs on until
 * the process ends, instead of from the handle's buffers; the result is
 * bit-identical to mpir_gmres_ldl_precond_dev's. A factor whose solves need
 * more than 64 MiB of work memory is rejected with MPIR_ERR_INVALID (the
 * message gives the bytes needed). The handle's device must be current.
 * Calls on one device share the array: they run one at a time (a lock held
 * until the result is complete), and any other use of the array overlapping
 * them would race. Single-system L D L^T handles only. */
int mpir_gmres_ldl_precond_dev_noalloc(mpir_gmres_ldl_t h, double* x);
/* average ms of `reps` runs measured with HIP events on the handle's stream:
 * what = 0 SpMV, 1 M^{-1} application, 2 numeric factorization, 3 M^{-1}
 * applied at once to as many vectors as the last batch solve had
 * right-hand sides (up to 8; 1 for a single system) */
int mpir_gmres_ldl_time(mpir_gmres_ldl_t h, int what, int reps, double* avg_ms);
/* One factorization of the current values with a timeline of the
 * dependency-driven launch (no counterpart in the reference): per task, in
 * list order, 4 entries: kind << 28 | supernode; the device wall clock
 * (100 MHz) when a workgroup drew it, when its waits ended, and when it was
 * counted (shifted left by 16, the workgroup index in the low 16 bits).
 * Writes at most cap / 4 tasks; *ntask receives the number of tasks. */
int mpir_gmres_ldl_trace_factor(mpir_gmres_ldl_t h, uint64_t* out, int64_t cap, int64_t* ntask);
/* One M^{-1} application (dependency-driven schedule) on a fixed vector with
 * the timeline of its tasks (no counterpart in the reference): per task, the
 * forward list then the backward list, 4 entries: kind | first supernode << 4;
 * the device wall clock (100 MHz) when drawn, when its wait ended, when done
 * (<< 16 | workgroup). Writes at most cap / 4 tasks; *ntask and *nfwd receive
 * the number of tasks and of forward tasks. */
int mpir_gmres_ldl_trace_precond(mpir_gmres_ldl_t h, uint64_t* out, int64_t cap, int64_t* ntask, int64_t* nfwd);
/* Per supernode of the analyzed plan, in the indices the two traces record
 * (no counterpart in the reference): its level in the supernodal tree (height
 * above the leaves), its columns and its rows below the diagonal block.
 * Writes them when cap >= the number of supernodes, which *ns receives. */
int mpir_gmres_ldl_supernode_info(mpir_gmres_ldl_t h, int32_t* level, int32_t* ncols, int32_t* nrows, int32_t cap,
                                  int32_t* ns);
/* Autotuner of the level-tile schedules (no counterpart in the reference,
 * whose solver has no device schedule). On a factored L D L^T handle (single
 * system or batch) it times, with HIP events over `reps` runs (1..100) each:
 * the factor's update grouping (panel steps whose trailing update is
 * applied at once: 2, 4, 8) x tile super-block edge (4, 8, 16 tiles), then
 * the triangular solves' narrow-front row limit (0, 256 .. 4096 rows) x
 * schedule (MPIR_SCHED_LEVELS / MPIR_SCHED_DAG), M^{-1} applied to as many
 * vectors at once as the handle's last batch solve (1 for a single
 * system); it keeps the fastest of
 * each, recomputes the factor with it, and later analyses of the handle
 * (same pattern or not) plan with the choice. Solutions stay within the
 * solver's tolerance; bits may differ from another tuning's, since sums are
 * taken in another order. choice (5, may be NULL): update grouping,
 * super-block edge, narrow-front rows, schedule, candidates timed;
 * ms (4, may be NULL): factor and M^{-1} ms with the choice, then with the
 * tuning in place before the call. */
int mpir_gmres_ldl_autotune(mpir_gmres_ldl_t h, int reps, int32_t* choice, double* ms);
/* Sets a tuning without timing it (the first three entries of autotune's
 * choice; no counterpart in the reference): update grouping 1..64,
 * super-block edge 1..64, narrow-front rows >= 0 (0: no narrow fronts). The
 * handle's plan is re-planned at once; an existing factor stays valid. The
 * defaults are 4, 8, 1024. */
int mpir_gmres_ldl_set_tuning(mpir_gmres_ldl_t h, int32_t upd_group, int32_t super_tiles, int32_t narrow_rows);
/* The log of the last solve in the JSON schema of the reference's VERBOSE
 * build (gmres_ir.hpp:119-124, 305-314, 539-631): max_outer_iterations,
 * max_inner_iterations, tolerance and an "events" array with one
 * "gmres_ir" event (initial_residual; per iteration: iteration, residual,
 * d_norm, timestamp_ms, inner_time_ms; stop_reason, stop_reason_detail,
 * total_iterations, total_time_ms). Writes at most cap bytes including the
 * terminating NUL and returns the full length (or a negative status). */
int64_t mpir_gmres_ldl_log(mpir_gmres_ldl_t h, char* buf, int64_t cap);
/* HIP stream of the handle (hipStream_t as void*) */
void* mpir_gmres_ldl_stream(mpir_gmres_ldl_t h);

/* ------------------------------------------------------------------ batches
 * Many independent systems on one handle (no counterpart in the reference,
 * whose GmresLDLIR holds one system; each system here is solved exactly as
 * GmresLDLIR::Solve, gmres_ir.hpp:534-741, would solve it). The systems are
 * joined into one block-diagonal system: one factorization plan over the
 * forest of their elimination trees and lock-step GMRES-IR with per-system,
 * per-right-hand-side scalars and stopping rules. A system's result is
 * bit-identical alone and in any batch.
 * The symbolic phase of each pattern is kept in a process-wide cache keyed by
 * (n, Ap, Ai, ordering), so repeated patterns are analysed once; missing
 * patterns are analysed on `threads` host threads. The handle's setters
 * (tolerance, iteration limits, ordering, scaling) apply to every system. */
int mpir_batch_analyze(mpir_gmres_ldl_t h, int32_t nsys, const int32_t* n, const int32_t* const* Ap,
                       const int32_t* const* Ai, int32_t threads);
/* values of every system, in the entry order of its pattern */
int mpir_batch_factor(mpir_gmres_ldl_t h, const double* const* Ax);
int mpir_batch_factor_dev(mpir_gmres_ldl_t h, const double* const* Ax);
/* nrhs[s] right-hand sides of system s, column-major with ld = n_s */
int mpir_batch_solve(mpir_gmres_ldl_t h, const int32_t* nrhs, const double* const* b, double* const* x);
int mpir_batch_solve_dev(mpir_gmres_ldl_t h, const int32_t* nrhs, const double* const* b, double* const* x);
/* results of the last batch: positive pivots per system (nsys); per system s
 * and right-hand side r at s * max_rhs + r: IR iterations, GMRES iterations,
 * stop reason (as mpir_gmres_stats, plus 4: the right-hand side or the
 * residual is not finite; the other instances of the batch go on), final
 * relative residual (scaled system), ms from the start of the solve until
 * the instance stopped. Any output pointer may be NULL. */
int mpir_batch_stats(mpir_gmres_ldl_t h, int32_t* max_rhs, int32_t* positive_pivots, int32_t* ir_iterations,
                     int32_t* gmres_iterations, int32_t* stop_reason, double* final_rel_residual, double* done_ms);
/* The batched solve runs every refinement iteration and Arnoldi step on the
 * device (the Hessenberg columns, Givens rotations, back-substitutions and
 * stopping rules of all instances in kernels), as one fixed launch sequence
 * up to the iteration limits in which each kernel returns at once when no
 * instance takes part in its iteration or step. When a solve repeats the
 * shapes, tolerance, plan and buffers of the solve before it, the sequence
 * is captured as a HIP graph and each later solve of that kind is one graph
 * launch (no counterpart in the reference). enable: 1 on (default), 0 off
 * (every solve launched kernel by kernel), -1 leave as is. counts (3, may
 * be NULL) of the last solve: graph nodes replayed (0: launched kernel by
 * kernel), Arnoldi steps and M^{-1} sweeps of the batch. */
int mpir_batch_graph(mpir_gmres_ldl_t h, int enable, int32_t* counts);
/* Plans the calling thread's QDLDL_solve engines have built from factor
 * patterns. A mpir_qdldl_solve_* call whose (Lp, Li) equals the previous
 * call's on the same thread (same factor and x precision) reuses its plan
 * and uploads only Lx and Dinv; no counterpart in the reference. */
int mpir_qdldl_solve_plan_builds(int64_t* builds);
/* the process-wide pattern cache: entries, hits, misses */
int mpir_pattern_cache_info(int64_t* entries, int64_t* hits, int64_t* misses);
/* Host only: symbolic phase of a batch without a device. info[0..7] = rows,
 * nnz(L), factor flops, supernodes, levels, panel entries, panel tasks,
 * forward solve tasks of the joined plan; info[8 + 2 s], info[9 + 2 s] =
 * nnz(L) and device-byte estimate of system s alone. */
int mpir_batch_plan_info(int32_t nsys, const int32_t* n, const int32_t* const* Ap, const int32_t* const* Ai,
                         int ordering, int64_t* info);

/* ------------------------------------------------------------------ serving
 * A stream of solve requests with continuous batching (no counterpart in the
 * reference, which solves one system per GmresLDLIR object). Each request
 * names a registered pattern and carries fresh values and nrhs right-hand
 * sides; its data are derived deterministically from (pattern, seed) by
 * mpir_server_request_data: off-diagonals of H and entries of A times
 * (1 + 0.05u), -D times (1 + 0.05u), diag(H) times (1.05 + 0.05v), u ~ U(-1,1),
 * v ~ U(0,1); x_true ~ U(-1,1); b = K x_true. Requests are admitted in arrival
 * order into batches while the device bytes of all batches in flight stay
 * within budget_bytes (a request larger than the budget runs alone);
 * `lanes` batches (host thread + HIP stream each) run concurrently, so
 * requests arriving while a batch is on the device join the next one and
 * overlap with it. Every request is solved as GmresLDLIR::Solve would solve
 * it (FP32 or FP64 factor, FP64 GMRES-IR to tol); its result is
 * bit-identical whichever batch it lands in. */
typedef struct mpir_server* mpir_server_t;
int mpir_server_create(int uf, int lanes, int64_t budget_bytes, int32_t max_batch, double tol, int ordering,
                       int32_t keep_x, int32_t threads, mpir_server_t* out);
void mpir_server_destroy(mpir_server_t s);
/* symbolic phase of the pattern (cached); base values and the size n1 of the
 * H block; returns the pattern id (>= 0) or a negative status */
int mpir_server_add_pattern(mpir_server_t s, int32_t n, const int32_t* Ap, const int32_t* Ai, const double* base,
                            int32_t n1);
/* `count` patterns at once, analysed on `threads` host threads; returns the
 * id of the first (the others follow in order) */
int mpir_server_add_patterns(mpir_server_t s, int32_t count, const int32_t* n, const int32_t* const* Ap,
                             const int32_t* const* Ai, const double* const* base, const int32_t* n1, int32_t threads);
/* Runs nreq requests arriving at t_arrive[i] ms after the call (ascending) to
 * completion. Outputs per request (any may be NULL): admission and
 * completion times (ms), batch and lane, status, the largest IR / GMRES
 * iteration count and final relative residual over its right-hand sides. */
int mpir_server_run(mpir_server_t s, int64_t nreq, const int32_t* pattern, const uint64_t* seed, const int32_t* nrhs,
                    const double* t_arrive, double* t_admit, double* t_done, int32_t* batch, int32_t* lane,
                    int32_t* status, int32_t* ir_max, int32_t* gmres_max, double* relres_max);
/* Online use: a request with the caller's values Ax (nnz of the pattern, in
 * its upper-CSC order) and nrhs right-hand sides b (n x nrhs, column-major),
 * both copied before the call returns; *id names it. The request joins the
 * queue at once and a free lane admits it into its next batch (the
 * reference's Compute + Solve of one system, gmres_ir.hpp:271-285). */
int mpir_server_submit(mpir_server_t s, int32_t pattern, const double* Ax, const double* b, int32_t nrhs,
                       int64_t* id);
typedef struct mpir_server_request {
  int32_t status;           /* 0, or the negative status of its batch */
  int32_t batch, lane;
  int32_t ir_max, gmres_max; /* over its right-hand sides */
  double rel_residual_max;   /* ||r||_inf / ||b_s||_inf, scaled system */
  double queue_ms;           /* submission to admission */
  double latency_ms;         /* submission to completion */
} mpir_server_request;
/* Waits up to timeout_ms (0: just look; < 0: no limit) for request id.
 * Returns 1 when it has completed (x, n x nrhs, and info written when
 * non-null), 0 while it is pending, or a negative status: the request
 * failed (info written, x not) or the call is invalid. With release != 0 a
 * completed request is forgotten and its id becomes invalid. Threads may
 * submit and poll concurrently. */
int mpir_server_poll(mpir_server_t s, int64_t id, double timeout_ms, double* x, mpir_server_request* info,
                     int32_t release);
/* counts[0..5] = completed, failed, batches, largest backlog (arrived, not
 * admitted), peak bytes in flight, most batches in flight at once;
 * ms[0..4] = wall time of the run, then summed over batches: request data,
 * joined symbolic plan and upload, numeric factorization, GMRES-IR */
int mpir_server_summary(mpir_server_t s, int64_t* counts, double* ms);
/* the two parts of the summary's plan time, summed over batches: joining
 * the systems' plans on the host, and the device copies, allocations and
 * the wait for them; and how often a device buffer of the process had to
 * grow during the run (each such reallocation waits for the whole device)
 * (no counterpart in the reference) */
int mpir_server_plan_phases(mpir_server_t s, double* join_ms, double* upload_ms, int64_t* regrowths);
/* solution of request i of the last run (keep_x != 0): n x nrhs, column-major */
int mpir_server_solution(mpir_server_t s, int64_t i, double* x);
/* the data of a request, on the host (Ax: nnz, b and x_true: n x nrhs) */
int mpir_server_request_data(mpir_server_t s, int32_t pattern, uint64_t seed, int32_t nrhs, double* Ax, double* b,
                             double* x_true);

/* ------------------------------------------------------------------ utilities
 * FP64 symmetric SpMV from the upper CSC with the reference's rounding
 * (MatrixMultiply, include/MPIR/ops.hpp:25-45), on the GPU. */
int mpir_spmv_upper_f64(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x, double* y);
/* r = b - K x in FP64 (ResidualCSCSymmetric, include/qdldl-luir/residual.hpp:7-44) */
int mpir_residual_upper_f64(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x,
                            const double* b, double* r);

/* Synthetic quasi-definite KKT systems (deterministic, SplitMix64):
 * dim 2/3 grid Laplacian of side k plus h_shift*I, m constraint rows
 * (a_mode 0: cnt_lo..cnt_hi distinct uniform columns, a_mode 1: a uniform
 * column and its +x,+y,+z grid neighbours, a_mode 2: cnt_lo..cnt_hi entries,
 * a uniform column and its nearest in-grid neighbours in a fixed offset
 * order), D = d_value*I, x_true ~ U(-1,1),
 * b = K x_true. Query sizes with NULL arrays, then fill. */
int mpir_gen_kkt(int dim, int k, double h_shift, int64_t m, int a_mode, int cnt_lo, int cnt_hi, double d_value,
                 uint64_t seed, int32_t* n_out, int64_t* nnz_out, int32_t* n1_out, int32_t* Ap, int32_t* Ai,
                 double* Ax, double* x_true, double* b);

/* AMD ordering of an upper CSC pattern (host): perm[k] = original index */
int mpir_amd_order(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* perm);
/* A Matrix Market coordinate file read as examples/example.cpp:87-138 reads
 * it: triplets transposed, sorted by (column, row), compressed to CSC (the
 * upper triangle of a symmetric file stored by its lower triangle). Call
 * with Ap = Ai = Ax = NULL for n / nnz, then with arrays of n+1 / nnz / nnz.
 * Host only. */
int mpir_read_mtx_upper(const char* path, int32_t* n, int64_t* nnz, int32_t* symmetric, int32_t* Ap, int32_t* Ai,
                        double* Ax);

#ifdef __cplusplus
}
#endif

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#endif /* MPIR_HIP_H */
