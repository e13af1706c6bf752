/*
 * ORACLE -- test infrastructure only. CPU restatement of the yex33/MPIR
 * solve path, used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker. The GPU product never calls this code.
 *
 * Each function restates the reference function named in its comment, in
 * plain C, in the same operation order (so FP64/FP32 results match the
 * reference bit for bit when compiled without contraction, -ffp-contract=off).
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   or_etree          - compiled reference headers (oracle/_ref), brute-force
 *                       symbolic elimination on tiny matrices
 *   or_factor_*       - oracle/_ref bit-exact; reconstruction ||K - LDL^T||
 *   or_solve_*        - oracle/_ref bit-exact; dense solve on tiny inputs
 *   or_matmul_upper   - oracle/_ref bit-exact; exact on integer inputs
 *   or_dnrm2/infnrm   - ops.hpp doctest known answers; oracle/_ref, also on
 *                       NaN / Inf / overflowing vectors
 *   or_scaling        - known answers of the |a_ii| > eps(UF) threshold for
 *                       each factor precision (tests/solve_rules.py)
 *   or_luir_*         - LUIR_QDLDL (include/qdldl-luir/luir.hpp:33-241) on the
 *                       pinned factor / solve / residual above; the loop
 *                       itself is pinned by closed forms (an exact FP64
 *                       factor stops after one correction; K = I) and by
 *                       outcome (luir.hpp does not build here: it includes
 *                       the absent QD library)
 *   or_gmres_ir       - closed forms (K = I, exact FP64 factor: one step with
 *                       happy breakdown), Arnoldi invariants, backward error;
 *                       the control layer by hand-derived traces: the
 *                       ||b||_2 < tol early exit, the rho > tol loop test,
 *                       the <= small-update stop, the absolute stagnation
 *                       stop, the clamp of the GMRES limit to n
 *                       (tests/test_oracle_solve_rules.py). Unpinned: that
 *                       the small-update test reads ||x|| before the update
 *                       (no input separates the two readings beyond a 10 eps
 *                       band), and the iteration-level trajectory of the
 *                       reference's LDL^T GmresLDLIR (the shipped class is
 *                       ILU-preconditioned and gmres_ir.hpp does not build
 *                       here: it needs <print>).
 */
#define _POSIX_C_SOURCE 199309L
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OR_UNKNOWN (-1)

/* QDLDL_etree, include/MPIR/qdldl.hpp:48-105 */
int64_t or_etree(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* work, int32_t* Lnz, int32_t* etree) {
  for (int32_t i = 0; i < n; i++) {
    work[i] = 0;
    Lnz[i] = 0;
    etree[i] = OR_UNKNOWN;
    if (Ap[i] == Ap[i + 1]) return -1;
  }
  for (int32_t j = 0; j < n; j++) {
    work[j] = j;
    for (int32_t p = Ap[j]; p < Ap[j + 1]; p++) {
      int32_t i = Ai[p];
      if (i > j) return -1;
      while (work[i] != j) {
        if (etree[i] == OR_UNKNOWN) etree[i] = j;
        Lnz[i]++;
        work[i] = j;
        i = etree[i];
      }
    }
  }
  int64_t sum = 0;
  for (int32_t i = 0; i < n; i++) {
    if (sum > INT32_MAX - Lnz[i]) return -2;
    sum += Lnz[i];
  }
  return sum;
}

/* QDLDL_factor + mul_sub, include/MPIR/qdldl.hpp:145-345 (mul_sub's 8-wide
 * blocking rounds each product before the subtraction exactly like the plain
 * loop, so the plain loop is used). */
#define OR_FACTOR(NAME, T)                                                                                          \
  int32_t NAME(int32_t n, const int32_t* Ap, const int32_t* Ai, const T* Ax, int32_t* Lp, int32_t* Li, T* Lx, T* D, \
               T* Dinv, const int32_t* Lnz, const int32_t* etree) {                                                  \
    unsigned char* yMarkers = (unsigned char*)calloc((size_t)n, 1);                                                  \
    int32_t* yIdx = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);                                                   \
    int32_t* elimBuffer = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);                                             \
    int32_t* LNext = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);                                                  \
    T* yVals = (T*)malloc(sizeof(T) * (size_t)n);                                                                    \
    int32_t positive = 0, ret;                                                                                       \
    Lp[0] = 0;                                                                                                       \
    for (int32_t i = 0; i < n; i++) {                                                                                \
      Lp[i + 1] = Lp[i] + Lnz[i];                                                                                    \
      yMarkers[i] = 0;                                                                                               \
      yVals[i] = (T)0.0;                                                                                             \
      D[i] = (T)0.0;                                                                                                 \
      LNext[i] = Lp[i];                                                                                              \
    }                                                                                                                \
    D[0] = Ax[0];                                                                                                    \
    if (D[0] == (T)0.0) { ret = -1; goto done; }                                                                     \
    if (D[0] > (T)0.0) positive++;                                                                                   \
    Dinv[0] = (T)1.0 / D[0];                                                                                         \
    for (int32_t k = 1; k < n; k++) {                                                                                \
      int32_t nnzY = 0;                                                                                              \
      for (int32_t i = Ap[k]; i < Ap[k + 1]; i++) {                                                                  \
        int32_t bidx = Ai[i];                                                                                        \
        if (bidx == k) { D[k] = Ax[i]; continue; }                                                                   \
        yVals[bidx] = Ax[i];                                                                                         \
        int32_t next = bidx;                                                                                         \
        if (!yMarkers[next]) {                                                                                       \
          yMarkers[next] = 1;                                                                                        \
          elimBuffer[0] = next;                                                                                      \
          int32_t nnzE = 1;                                                                                          \
          next = etree[bidx];                                                                                        \
          while (next != OR_UNKNOWN && next < k) {                                                                   \
            if (yMarkers[next]) break;                                                                               \
            yMarkers[next] = 1;                                                                                      \
            elimBuffer[nnzE++] = next;                                                                               \
            next = etree[next];                                                                                      \
          }                                                                                                          \
          while (nnzE) yIdx[nnzY++] = elimBuffer[--nnzE];                                                            \
        }                                                                                                            \
      }                                                                                                              \
      for (int32_t i = nnzY; i-- > 0;) {                                                                             \
        int32_t cidx = yIdx[i];                                                                                      \
        int32_t tmpIdx = LNext[cidx];                                                                                \
        T yc = yVals[cidx];                                                                                          \
        for (int32_t j = Lp[cidx]; j < tmpIdx; j++) yVals[Li[j]] -= Lx[j] * yc;                                      \
        Li[tmpIdx] = k;                                                                                              \
        Lx[tmpIdx] = yc * Dinv[cidx];                                                                                \
        D[k] -= yc * Lx[tmpIdx];                                                                                     \
        LNext[cidx]++;                                                                                               \
        yVals[cidx] = (T)0.0;                                                                                        \
        yMarkers[cidx] = 0;                                                                                          \
      }                                                                                                              \
      if (D[k] == (T)0.0) { ret = -1; goto done; }                                                                   \
      if (D[k] > (T)0.0) positive++;                                                                                 \
      Dinv[k] = (T)1.0 / D[k];                                                                                       \
    }                                                                                                                \
    ret = positive;                                                                                                  \
  done:                                                                                                              \
    free(yMarkers); free(yIdx); free(elimBuffer); free(LNext); free(yVals);                                          \
    return ret;                                                                                                      \
  }
OR_FACTOR(or_factor_f64, double)
OR_FACTOR(or_factor_f32, float)

/* QDLDL_Lsolve / Ltsolve / solve, include/MPIR/qdldl.hpp:362-432 */
#define OR_SOLVE(NAME, TF, TX)                                                                       \
  void NAME(int32_t n, const int32_t* Lp, const int32_t* Li, const TF* Lx, const TF* Dinv, TX* x) {  \
    for (int32_t i = 0; i < n; i++) {                                                                \
      TX val = x[i];                                                                                 \
      for (int32_t j = Lp[i]; j < Lp[i + 1]; j++) x[Li[j]] -= (TX)Lx[j] * val;                       \
    }                                                                                                \
    for (int32_t i = 0; i < n; i++) x[i] *= (TX)Dinv[i];                                             \
    for (int32_t i = n; i-- > 0;) {                                                                  \
      TX val = x[i];                                                                                 \
      for (int32_t j = Lp[i]; j < Lp[i + 1]; j++) val -= (TX)Lx[j] * x[Li[j]];                       \
      x[i] = val;                                                                                    \
    }                                                                                                \
  }
OR_SOLVE(or_solve_f64_f64, double, double)
OR_SOLVE(or_solve_f32_f64, float, double)
OR_SOLVE(or_solve_f32_f32, float, float)

/* MatrixMultiply<double>, include/MPIR/ops.hpp:25-45 (symmetric, upper CSC) */
void or_matmul_upper(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x, double* res) {
  for (int32_t i = 0; i < n; i++) res[i] = 0.0;
  for (int32_t col = 0; col < n; col++)
    for (int32_t idx = Ap[col]; idx < Ap[col + 1]; idx++) {
      const int32_t row = Ai[idx];
      res[col] += Ax[idx] * x[row];
      if (row < col) res[row] += Ax[idx] * x[col];
    }
}

/* Dnrm2<double>, include/MPIR/ops.hpp:231-255 (scaled sum of squares, last to first) */
double or_dnrm2(int64_t n, const double* x) {
  if (n == 0) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int64_t i = n; i-- > 0;) {
    if (x[i] != 0.0) {
      const double a = fabs(x[i]);
      if (scale < a) {
        ssq = 1.0 + ssq * ((scale / a) * (scale / a));
        scale = a;
      } else {
        ssq = ssq + ((a / scale) * (a / scale));
      }
    }
  }
  return scale * sqrt(ssq);
}

/* VectorDot<double>, include/MPIR/ops.hpp:195-200: std::transform_reduce
 * with std::plus / std::multiplies. libstdc++ (GCC 11.4.0, the toolchain
 * here) evaluates it for random-access iterators four terms at a time:
 * init += ((a0 b0 + a1 b1) + (a2 b2 + a3 b3)), then the tail one by one. */
double or_dot(int64_t n, const double* a, const double* b) {
  double init = 0.0;
  int64_t i = 0;
  for (; n - i >= 4; i += 4) {
    const double v1 = a[i] * b[i] + a[i + 1] * b[i + 1];
    const double v2 = a[i + 2] * b[i + 2] + a[i + 3] * b[i + 3];
    init = init + (v1 + v2);
  }
  for (; i < n; i++) init = init + a[i] * b[i];
  return init;
}

/* InfNrm, include/MPIR/ops.hpp:324-329: |*std::max_element(x, |a| < |b|)|.
 * max_element keeps the first element and replaces it only when
 * |kept| < |x[i]| holds, so a NaN in x[0] is returned (every comparison with
 * it is false) and a NaN anywhere else is skipped. The reference
 * dereferences end() for an empty vector; this returns 0 there. */
double or_infnrm(int64_t n, const double* x) {
  if (n <= 0) return 0.0;
  double kept = x[0];
  for (int64_t i = 1; i < n; i++)
    if (fabs(kept) < fabs(x[i])) kept = x[i];
  return fabs(kept);
}

/* ------------------------------------------------------------------ GMRES-IR */
typedef struct or_ilu or_ilu;
static void or_ilu_solve(const or_ilu* f, double* x);

typedef struct {
  int32_t n;
  const int32_t *Ap, *Ai;
  const double* Ax; /* scaled values, original order (UW) */
  const or_ilu* ilu; /* the FGPILU factors as M (instead of LDL^T), or NULL */
  /* factor of the permuted scaled matrix */
  const int32_t* perm; /* new -> old, or NULL */
  int32_t uf;          /* 0 double, 1 float */
  int32_t *Lp, *Li;
  void *Lx, *Dinv;
  double* work; /* n */
  int64_t gmres_iter;
  double tol;
  int64_t gmres_total;
  int64_t precond_count;
} or_ctx;

/* M^{-1} v: QDLDL_solve<INT, UF, UW> on the permuted system */
static void or_precond(or_ctx* c, double* v) {
  const int32_t n = c->n;
  c->precond_count++;
  if (c->ilu) {
    or_ilu_solve(c->ilu, v);
    return;
  }
  if (c->perm) {
    for (int32_t k = 0; k < n; k++) c->work[k] = v[c->perm[k]];
  } else {
    memcpy(c->work, v, sizeof(double) * (size_t)n);
  }
  if (c->uf == 0) or_solve_f64_f64(n, c->Lp, c->Li, (const double*)c->Lx, (const double*)c->Dinv, c->work);
  else or_solve_f32_f64(n, c->Lp, c->Li, (const float*)c->Lx, (const float*)c->Dinv, c->work);
  if (c->perm) {
    for (int32_t k = 0; k < n; k++) v[c->perm[k]] = c->work[k];
  } else {
    memcpy(v, c->work, sizeof(double) * (size_t)n);
  }
}

/* GmresLDLIR::PrecondGmres, include/MPIR/gmres_ir.hpp:636-741 (MGS with the
 * Brown/Hindmarsh reorthogonalisation, Givens rotations, back substitution) */
static void or_precond_gmres(or_ctx* c, const double* b, double* res) {
  const int32_t n = c->n;
  if (c->gmres_iter > n) c->gmres_iter = n;
  const int64_t m = c->gmres_iter;
  for (int32_t i = 0; i < n; i++) res[i] = 0.0;
  double rho = or_dnrm2(n, b);
  if (rho < c->tol) return;
  double* V = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(m + 1));
  double* H = (double*)calloc((size_t)(m + 1) * (size_t)(m > 0 ? m : 1), sizeof(double)); /* H[i*m + j] */
  double* g = (double*)calloc((size_t)m + 1, sizeof(double));
  double* cs = (double*)calloc((size_t)m + 1, sizeof(double));
  double* sn = (double*)calloc((size_t)m + 1, sizeof(double));
  memcpy(V, b, sizeof(double) * (size_t)n);
  or_precond(c, V);
  rho = or_dnrm2(n, V);
  const double inv0 = 1 / rho;
  for (int32_t i = 0; i < n; i++) V[i] = V[i] * inv0;
  g[0] = rho;
#define HH(i, j) H[(size_t)(i) * (size_t)m + (size_t)(j)]
  int64_t k;
  for (k = 0; k < m && rho > c->tol; k++) {
    double* w = V + (size_t)n * (size_t)(k + 1);
    or_matmul_upper(n, c->Ap, c->Ai, c->Ax, V + (size_t)n * (size_t)k, w);
    or_precond(c, w);
    const double normav = or_dnrm2(n, w);
    for (int64_t j = 0; j <= k; j++) {
      const double* vj = V + (size_t)n * (size_t)j;
      const double dot = or_dot(n, w, vj);
      HH(j, k) = dot;
      for (int32_t i = 0; i < n; i++) w[i] = w[i] - vj[i] * dot;
    }
    const double normav2 = or_dnrm2(n, w);
    HH(k + 1, k) = normav2;
    if (normav + 0.001 * normav2 == normav) {
      for (int64_t j = 0; j <= k; j++) {
        const double* vj = V + (size_t)n * (size_t)j;
        const double hr = or_dot(n, w, vj);
        HH(j, k) += hr;
        for (int32_t i = 0; i < n; i++) w[i] = w[i] - vj[i] * hr;
      }
      HH(k + 1, k) = or_dnrm2(n, w);
    }
    if (HH(k + 1, k) != 0) {
      const double inv = 1 / HH(k + 1, k);
      for (int32_t i = 0; i < n; i++) w[i] = w[i] * inv;
    }
    for (int64_t j = 0; j < k; j++) {
      const double w1 = cs[j] * HH(j, k) - sn[j] * HH(j + 1, k);
      const double w2 = sn[j] * HH(j, k) + cs[j] * HH(j + 1, k);
      HH(j, k) = w1;
      HH(j + 1, k) = w2;
    }
    const double hv[2] = {HH(k, k), HH(k + 1, k)};
    const double nu = or_dnrm2(2, hv);
    if (nu != 0) {
      cs[k] = HH(k, k) / nu;
      sn[k] = -HH(k + 1, k) / nu;
      HH(k, k) = cs[k] * HH(k, k) - sn[k] * HH(k + 1, k);
      HH(k + 1, k) = 0.0;
      const double w1 = cs[k] * g[k] - sn[k] * g[k + 1];
      const double w2 = sn[k] * g[k] + cs[k] * g[k + 1];
      g[k] = w1;
      g[k + 1] = w2;
    }
    rho = fabs(g[k + 1]);
  }
  c->gmres_total += k;
  double* y = (double*)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
  for (int64_t i = k; i-- > 0;) {
    y[i] = g[i];
    for (int64_t j = i + 1; j < k; j++) y[i] -= HH(i, j) * y[j];
    y[i] /= HH(i, i);
  }
  for (int32_t i = 0; i < n; i++)
    for (int64_t j = 0; j < k; j++) res[i] += V[(size_t)n * (size_t)j + (size_t)i] * y[j];
#undef HH
  free(y); free(V); free(H); free(g); free(cs); free(sn);
}

/* permute the scaled upper CSC: C = triu(P A P^T), rows ascending */
static void or_permute(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const int32_t* perm,
                       int32_t* Cp, int32_t* Ci, double* Cx) {
  int32_t* pinv = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* cnt = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  for (int32_t k = 0; k < n; k++) pinv[perm[k]] = k;
  for (int32_t j = 0; j < n; j++)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; q++) {
      const int32_t a = pinv[Ai[q]], b = pinv[j];
      cnt[(a > b ? a : b) + 1]++;
    }
  Cp[0] = 0;
  for (int32_t j = 0; j < n; j++) Cp[j + 1] = Cp[j] + cnt[j + 1];
  for (int32_t j = 0; j < n; j++) cnt[j] = Cp[j];
  for (int32_t j = 0; j < n; j++)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; q++) {
      const int32_t a = pinv[Ai[q]], b = pinv[j];
      const int32_t col = a > b ? a : b, row = a > b ? b : a;
      const int32_t pos = cnt[col]++;
      Ci[pos] = row;
      Cx[pos] = Ax[q];
    }
  for (int32_t j = 0; j < n; j++) /* insertion sort per column */
    for (int32_t q = Cp[j] + 1; q < Cp[j + 1]; q++) {
      const int32_t r = Ci[q];
      const double v = Cx[q];
      int32_t t = q - 1;
      while (t >= Cp[j] && Ci[t] > r) {
        Ci[t + 1] = Ci[t];
        Cx[t + 1] = Cx[t];
        t--;
      }
      Ci[t + 1] = r;
      Cx[t + 1] = v;
    }
  free(pinv);
  free(cnt);
}

/*
 * GmresLDLIR<UF, double, double>: Compute (gmres_ir.hpp:323-363 scaling, then
 * QDLDL_etree + QDLDL_factor<UF> of the MIS, MIS.tex:505-513, here of
 * P K_s P^T when perm is given) and Solve (gmres_ir.hpp:534-634).
 * info[0..7] = ir iterations, gmres iterations, stop reason (0 max, 1 small
 * update, 2 stagnation, 3 exact zero residual), factor return, precond applications, final r_norm,
 * factor seconds, solve seconds.
 */
int32_t or_gmres_ir(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const int32_t* perm, int32_t uf,
                    int64_t ir_iter, int64_t gmres_iter, double tol, const double* b, double* x, double* info);
void or_scaling(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, int32_t uf, double* AD,
                double* Axs);


/* LUIR_QDLDL<FACTPREC, double, double, double, int32_t>,
 * include/qdldl-luir/luir.hpp:33-241: QDLDL_etree + QDLDL_factor on the
 * matrix as given (values rounded to FACTPREC, :120-148); x0 = solve in
 * FACTPREC (:173-178); refinement (:196-225): correction = FACTPREC(b - A x)
 * with the residual in double (ResidualCSCSymmetric, residual.hpp:7-44),
 * scaled by 1/||.||_inf (tblas.hpp scale: x[i] *= s), solved in FACTPREC,
 * scaled back, x += correction in double; stop when
 * ||d||/||x|| <= 10 eps(double) or ||d||/||d_prev|| >= 0.5. maxiter = 0 is a
 * single solve in double (:162-170). One deviation: an exactly zero residual
 * ends the loop (x is exact) where the reference divides by its zero norm
 * (:203) and turns x into NaN. info: [0] refinement steps,
 * [1] ||b - A x||_inf / ||b||_inf (residual rounded to FACTPREC first),
 * [2] sum(Lnz). Returns QDLDL_factor's value (< 0 on failure). */
static void or_residual_upper(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x,
                              const double* b, double* r) {
  or_matmul_upper(n, Ap, Ai, Ax, x, r);
  for (int32_t i = 0; i < n; i++) r[i] = b[i] - r[i];
}

#define OR_LUIR(NAME, TF, FACTOR, SOLVE_TF, SOLVE_TX)                                                         \
  int32_t NAME(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b, double* x, \
               int32_t maxiter, double* info) {                                                              \
    int32_t* work = (int32_t*)malloc(sizeof(int32_t) * n);                                                    \
    int32_t* Lnz = (int32_t*)malloc(sizeof(int32_t) * n);                                                     \
    int32_t* etree = (int32_t*)malloc(sizeof(int32_t) * n);                                                   \
    int64_t sum = or_etree(n, Ap, Ai, work, Lnz, etree);                                                      \
    if (sum < 0) { free(work); free(Lnz); free(etree); return (int32_t)sum; }                                 \
    const int32_t nnz = Ap[n];                                                                                \
    int32_t* Lp = (int32_t*)malloc(sizeof(int32_t) * (n + 1));                                                \
    int32_t* Li = (int32_t*)malloc(sizeof(int32_t) * (sum > 0 ? sum : 1));                                   \
    TF* Lx = (TF*)malloc(sizeof(TF) * (sum > 0 ? sum : 1));                                                   \
    TF* D = (TF*)malloc(sizeof(TF) * n);                                                                      \
    TF* Dinv = (TF*)malloc(sizeof(TF) * n);                                                                   \
    TF* Af = (TF*)malloc(sizeof(TF) * (nnz > 0 ? nnz : 1));                                                   \
    for (int32_t q = 0; q < nnz; q++) Af[q] = (TF)Ax[q];                                                      \
    int32_t ret = FACTOR(n, Ap, Ai, Af, Lp, Li, Lx, D, Dinv, Lnz, etree);                                     \
    info[2] = (double)sum;                                                                                    \
    double* r = (double*)malloc(sizeof(double) * n);                                                          \
    TF* c = (TF*)malloc(sizeof(TF) * n);                                                                      \
    info[0] = 0;                                                                                              \
    if (ret >= 0) {                                                                                           \
      double bn = 0;                                                                                          \
      for (int32_t i = 0; i < n; i++) bn = fabs(b[i]) > bn ? fabs(b[i]) : bn;                                 \
      if (maxiter == 0) {                                                                                     \
        for (int32_t i = 0; i < n; i++) x[i] = b[i];                                                          \
        SOLVE_TX(n, Lp, Li, Lx, Dinv, x);                                                                     \
      } else {                                                                                                \
        for (int32_t i = 0; i < n; i++) c[i] = (TF)b[i];                                                      \
        SOLVE_TF(n, Lp, Li, Lx, Dinv, c);                                                                     \
        for (int32_t i = 0; i < n; i++) x[i] = (double)c[i];                                                  \
        double d0 = DBL_MAX;                                                                                  \
        for (int32_t it = 1; it <= maxiter; it++) {                                                           \
          or_residual_upper(n, Ap, Ai, Ax, x, b, r);                                                          \
          TF cm = 0;                                                                                          \
          for (int32_t i = 0; i < n; i++) {                                                                   \
            c[i] = (TF)r[i];                                                                                  \
            TF a = c[i] < 0 ? -c[i] : c[i];                                                                   \
            if (a > cm) cm = a;                                                                               \
          }                                                                                                   \
          const double nc = (double)cm;                                                                       \
          if (nc == 0.0) { info[0] = it; break; } /* deviation: the reference divides by zero here */         \
          for (int32_t i = 0; i < n; i++) c[i] = (TF)((double)c[i] * (1. / nc));                              \
          SOLVE_TF(n, Lp, Li, Lx, Dinv, c);                                                                   \
          TF dm = 0;                                                                                          \
          for (int32_t i = 0; i < n; i++) {                                                                   \
            c[i] = (TF)((double)c[i] * nc);                                                                   \
            TF a = c[i] < 0 ? -c[i] : c[i];                                                                   \
            if (a > dm) dm = a;                                                                               \
          }                                                                                                   \
          double xm = 0;                                                                                      \
          for (int32_t i = 0; i < n; i++) xm = fabs(x[i]) > xm ? fabs(x[i]) : xm;                             \
          const double dn = (double)dm;                                                                       \
          const int stop = dn / xm <= 10 * DBL_EPSILON || dn / d0 >= 0.5;                                     \
          info[0] = it;                                                                                       \
          for (int32_t i = 0; i < n; i++) x[i] = (double)c[i] + x[i];                                         \
          if (stop) break;                                                                                    \
          d0 = dn;                                                                                            \
        }                                                                                                     \
      }                                                                                                       \
      or_residual_upper(n, Ap, Ai, Ax, x, b, r);                                                              \
      TF rm = 0;                                                                                              \
      for (int32_t i = 0; i < n; i++) {                                                                       \
        TF a = (TF)r[i];                                                                                      \
        a = a < 0 ? -a : a;                                                                                   \
        if (a > rm) rm = a;                                                                                   \
      }                                                                                                       \
      info[1] = (double)rm / bn;                                                                              \
    }                                                                                                         \
    free(work); free(Lnz); free(etree); free(Lp); free(Li); free(Lx); free(D); free(Dinv); free(Af);          \
    free(r); free(c);                                                                                         \
    return ret;                                                                                               \
  }
OR_LUIR(or_luir_f64, double, or_factor_f64, or_solve_f64_f64, or_solve_f64_f64)
OR_LUIR(or_luir_f32, float, or_factor_f32, or_solve_f32_f32, or_solve_f32_f64)

static double or_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* GmresLDLIR::Compute's diagonal scaling, include/MPIR/gmres_ir.hpp:347-363:
 * AD[i] = 1/sqrt|a_ii| where |a_ii| > epsilon of the factor precision UF
 * (uf: 0 double, 1 float, 2 binary16, 3 bfloat16), else 1; then
 * a_ij *= AD[i] * AD[j]. Axs may be NULL (AD only). */
void or_scaling(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, int32_t uf, double* AD,
                double* Axs) {
  const double eps_uf = uf == 0   ? DBL_EPSILON
                        : uf == 1 ? (double)FLT_EPSILON
                        : uf == 2 ? 0.0009765625 /* 2^-10 */
                                  : 0.0078125;   /* 2^-7 */
  for (int32_t i = 0; i < n; i++) AD[i] = 1;
  for (int32_t col = 0; col < n; col++)
    for (int32_t idx = Ap[col]; idx < Ap[col + 1]; idx++) {
      const int32_t row = Ai[idx];
      if (row == col && fabs(Ax[idx]) > eps_uf) AD[row] = 1.0 / sqrt(fabs(Ax[idx]));
    }
  if (Axs)
    for (int32_t col = 0; col < n; col++)
      for (int32_t idx = Ap[col]; idx < Ap[col + 1]; idx++) Axs[idx] = Ax[idx] * (AD[Ai[idx]] * AD[col]);
}

int32_t or_gmres_ir(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const int32_t* perm, int32_t uf,
                    int64_t ir_iter, int64_t gmres_iter, double tol, const double* b, double* x, double* info) {
  const double t0 = or_now();
  const int32_t nnz = Ap[n];
  double* Axs = (double*)malloc(sizeof(double) * (size_t)nnz);
  double* AD = (double*)malloc(sizeof(double) * (size_t)n);
  or_scaling(n, Ap, Ai, Ax, uf, AD, Axs);

  int32_t *Cp = NULL, *Ci = NULL;
  double* Cx = NULL;
  const int32_t *Fp = Ap, *Fi = Ai;
  const double* Fx = Axs;
  if (perm) {
    Cp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    Ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
    Cx = (double*)malloc(sizeof(double) * (size_t)nnz);
    or_permute(n, Ap, Ai, Axs, perm, Cp, Ci, Cx);
    Fp = Cp;
    Fi = Ci;
    Fx = Cx;
  }
  int32_t* work = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* Lnz = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* etree = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  const int64_t sumLnz = or_etree(n, Fp, Fi, work, Lnz, etree);
  int32_t fret = -3;
  or_ctx c;
  memset(&c, 0, sizeof(c));
  double* wv = (double*)malloc(sizeof(double) * (size_t)n);
  if (sumLnz >= 0) {
    c.Lp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    c.Li = (int32_t*)malloc(sizeof(int32_t) * (size_t)(sumLnz > 0 ? sumLnz : 1));
    const size_t fs = uf == 0 ? sizeof(double) : sizeof(float);
    c.Lx = malloc(fs * (size_t)(sumLnz > 0 ? sumLnz : 1));
    c.Dinv = malloc(fs * (size_t)n);
    void* D = malloc(fs * (size_t)n);
    if (uf == 0) {
      fret = or_factor_f64(n, Fp, Fi, Fx, c.Lp, c.Li, (double*)c.Lx, (double*)D, (double*)c.Dinv, Lnz, etree);
    } else {
      float* fx = (float*)malloc(sizeof(float) * (size_t)nnz);
      for (int32_t q = 0; q < nnz; q++) fx[q] = (float)Fx[q];
      fret = or_factor_f32(n, Fp, Fi, fx, c.Lp, c.Li, (float*)c.Lx, (float*)D, (float*)c.Dinv, Lnz, etree);
      free(fx);
    }
    free(D);
  }
  const double t1 = or_now();
  int32_t rc = 0;
  int64_t iter = 0;
  int32_t stop = 0;
  double r_norm = 0.0;
  if (fret < 0) {
    rc = -1;
  } else {
    c.n = n;
    c.Ap = Ap;
    c.Ai = Ai;
    c.Ax = Axs;
    c.perm = perm;
    c.uf = uf;
    c.work = wv;
    c.gmres_iter = gmres_iter;
    c.tol = tol;
    double* bs = (double*)malloc(sizeof(double) * (size_t)n);
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* t = (double*)malloc(sizeof(double) * (size_t)n);
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    for (int32_t i = 0; i < n; i++) bs[i] = b[i] * AD[i];
    or_precond_gmres(&c, bs, x);
    or_matmul_upper(n, Ap, Ai, Axs, x, t);
    for (int32_t i = 0; i < n; i++) r[i] = bs[i] - t[i];
    const double bnorm = or_infnrm(n, bs);
    double r_prev = DBL_MAX;
    for (iter = 1; iter <= ir_iter; iter++) {
      const double scale = or_infnrm(n, r);
      /* deviation: an exactly zero residual would make the reference scale
       * by 1/0 (gmres_ir.hpp:562-563) and continue with NaN; stop instead */
      if (scale == 0.0) { stop = 3; break; }
      const double inv = 1.0 / scale;
      for (int32_t i = 0; i < n; i++) t[i] = r[i] * inv;
      or_precond_gmres(&c, t, d);
      for (int32_t i = 0; i < n; i++) d[i] = d[i] * scale;
      const double d_norm = or_infnrm(n, d);
      const double x_norm = or_infnrm(n, x);
      for (int32_t i = 0; i < n; i++) x[i] = x[i] + d[i];
      or_matmul_upper(n, Ap, Ai, Axs, x, t);
      for (int32_t i = 0; i < n; i++) r[i] = bs[i] - t[i];
      r_norm = or_infnrm(n, r) / bnorm;
      if (d_norm <= x_norm * (10 * DBL_EPSILON)) { stop = 1; break; }
      if (fabs(r_norm - r_prev) < tol) { stop = 2; break; }
      r_prev = r_norm;
    }
    if (iter > ir_iter) iter = ir_iter;
    for (int32_t i = 0; i < n; i++) x[i] = x[i] * AD[i];
    free(bs); free(r); free(t); free(d);
  }
  const double t2 = or_now();
  if (info) {
    info[0] = (double)iter;
    info[1] = (double)c.gmres_total;
    info[2] = (double)stop;
    info[3] = (double)fret;
    info[4] = (double)c.precond_count;
    info[5] = r_norm;
    info[6] = t1 - t0;
    info[7] = t2 - t1;
  }
  free(wv); free(work); free(Lnz); free(etree);
  free(c.Lp); free(c.Li); free(c.Lx); free(c.Dinv);
  free(Cp); free(Ci); free(Cx); free(Axs); free(AD);
  return rc;
}

/* ------------------------------------------------------------------ FGPILU
 * GmresLDLIR::Compute's ILU path, include/MPIR/gmres_ir.hpp:323-531, with
 * SparseLDotU (:776-799), NLResNorm (:801-824), ValueAt (:826-836) and the
 * M^{-1} slot TriangularSolve (:743-774), run as the single-threaded
 * OpenMP loop runs: columns ascending, entries ascending, updates in place.
 * U is CSC (Up, Ui), L is CSR with the same index arrays (row r of L holds
 * the entries of column r of U, mirrored). UF: the factor precision (float
 * or double); the nonlinear residual is summed in double. */
struct or_ilu {
  int32_t n, uf;
  int64_t nnz;
  int32_t *Up, *Ui;
  double* ax; /* A(Ui[idx], col) of each pattern entry (0 if absent) */
  void *Ux, *Lx, *UDinv;
  int32_t sweeps;
  double nl_res0, nl_res;
};

/* the ILU_k pattern: Compute's loop at :369-408 (the upper pattern of A,
 * then k times: column col collects the rows of A's columns row1 for every
 * row1 in its current pattern, sorted) */
static void or_ilu_pattern(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t ilu_k, int32_t** Up_out,
                           int32_t** Ui_out) {
  int32_t* Up = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  memcpy(Up, Ap, sizeof(int32_t) * ((size_t)n + 1));
  int32_t* Ui = (int32_t*)malloc(sizeof(int32_t) * (size_t)(Ap[n] > 0 ? Ap[n] : 1));
  memcpy(Ui, Ai, sizeof(int32_t) * (size_t)Ap[n]);
  char* marker = (char*)calloc((size_t)n, 1);
  int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int32_t it = 0; it < ilu_k; it++) {
    int32_t* Np = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    size_t cap = (size_t)Up[n] * 2 + 16, used = 0;
    int32_t* Ni = (int32_t*)malloc(sizeof(int32_t) * cap);
    Np[0] = 0;
    for (int32_t col = 0; col < n; col++) {
      int32_t cnt = 0;
      for (int32_t i1 = Up[col]; i1 < Up[col + 1]; i1++) {
        const int32_t row1 = Ui[i1];
        if (row1 > col) continue;
        for (int32_t i2 = Ap[row1]; i2 < Ap[row1 + 1]; i2++) {
          const int32_t row2 = Ai[i2];
          if (row2 > col) continue;
          if (!marker[row2]) {
            marker[row2] = 1;
            rows[cnt++] = row2;
          }
        }
      }
      /* sort (insertion sort: columns are short) and clear the marks */
      for (int32_t a = 1; a < cnt; a++) {
        const int32_t v = rows[a];
        int32_t b = a - 1;
        while (b >= 0 && rows[b] > v) {
          rows[b + 1] = rows[b];
          b--;
        }
        rows[b + 1] = v;
      }
      if (used + (size_t)cnt > cap) {
        cap = (used + (size_t)cnt) * 2;
        Ni = (int32_t*)realloc(Ni, sizeof(int32_t) * cap);
      }
      for (int32_t a = 0; a < cnt; a++) {
        Ni[used++] = rows[a];
        marker[rows[a]] = 0;
      }
      Np[col + 1] = (int32_t)used;
    }
    free(Up);
    free(Ui);
    Up = Np;
    Ui = Ni;
  }
  free(marker);
  free(rows);
  *Up_out = Up;
  *Ui_out = Ui;
}

#define OR_ILU_BODY(UF)                                                                                     \
  /* SparseLDotU: sum_k L[row, k] U[k, col] over k < cut (:776-799) */                                     \
  UF* Ux = (UF*)f->Ux;                                                                                      \
  UF* Lx = (UF*)f->Lx;                                                                                      \
  UF* UDinv = (UF*)f->UDinv;                                                                                \
  const int32_t *Up = f->Up, *Ui = f->Ui;                                                                   \
  const int32_t n = f->n;

static float or_ldotu_f32(const or_ilu* f, int32_t row, int32_t col, int32_t cut) {
  OR_ILU_BODY(float)
  (void)UDinv;
  (void)n;
  int32_t i1 = Up[row], i2 = Up[col];
  float sum = 0.0f;
  while (i1 < Up[row + 1] && i2 < Up[col + 1]) {
    const int32_t kL = Ui[i1], kU = Ui[i2];
    if (kL >= cut || kU >= cut) break;
    if (kL == kU) {
      sum += Lx[i1] * Ux[i2];
      i1++;
      i2++;
    } else if (kL < kU) {
      i1++;
    } else {
      i2++;
    }
  }
  return sum;
}
static double or_ldotu_f64(const or_ilu* f, int32_t row, int32_t col, int32_t cut) {
  OR_ILU_BODY(double)
  (void)UDinv;
  (void)n;
  int32_t i1 = Up[row], i2 = Up[col];
  double sum = 0.0;
  while (i1 < Up[row + 1] && i2 < Up[col + 1]) {
    const int32_t kL = Ui[i1], kU = Ui[i2];
    if (kL >= cut || kU >= cut) break;
    if (kL == kU) {
      sum += Lx[i1] * Ux[i2];
      i1++;
      i2++;
    } else if (kL < kU) {
      i1++;
    } else {
      i2++;
    }
  }
  return sum;
}

/* NLResNorm (:801-824): sum |A - L U| over the U entries, then over the
 * strictly-lower L entries (row r of L holds the entries of column r of U) */
static double or_nlres(const or_ilu* f) {
  double res = 0.0;
  const int32_t n = f->n;
  for (int32_t col = 0; col < n; col++)
    for (int32_t idx = f->Up[col]; idx < f->Up[col + 1]; idx++) {
      const int32_t row = f->Ui[idx];
      if (f->uf == 1) {
        const float ax = (float)f->ax[idx];
        const float sum = or_ldotu_f32(f, row, col, row + 1);
        res += fabsf(ax - sum);
      } else {
        const double ax = f->ax[idx];
        const double sum = or_ldotu_f64(f, row, col, row + 1);
        res += fabs(ax - sum);
      }
    }
  for (int32_t row = 0; row < n; row++)
    for (int32_t idx = f->Up[row]; idx < f->Up[row + 1]; idx++) {
      const int32_t col = f->Ui[idx];
      if (col >= row) continue;
      if (f->uf == 1) {
        const float ax = (float)f->ax[idx];
        const float sum = or_ldotu_f32(f, row, col, col + 1);
        res += fabsf(ax - sum);
      } else {
        const double ax = f->ax[idx];
        const double sum = or_ldotu_f64(f, row, col, col + 1);
        res += fabs(ax - sum);
      }
    }
  return res;
}

/* initial guesses (:414-441) and the FGPILU sweeps (:447-518) */
static void or_fgpilu_run(or_ilu* f, int32_t max_sweeps, double tol) {
  const int32_t n = f->n;
  for (int32_t col = 0; col < n; col++)
    for (int32_t idx = f->Up[col]; idx < f->Up[col + 1]; idx++) {
      const int32_t row = f->Ui[idx];
      if (f->uf == 1) {
        ((float*)f->Ux)[idx] = (float)f->ax[idx];
        if (row == col) ((float*)f->UDinv)[col] = 1.0f / ((float*)f->Ux)[idx];
      } else {
        ((double*)f->Ux)[idx] = f->ax[idx];
        if (row == col) ((double*)f->UDinv)[col] = 1.0 / ((double*)f->Ux)[idx];
      }
    }
  memcpy(f->Lx, f->Ux, (f->uf == 1 ? sizeof(float) : sizeof(double)) * (size_t)f->nnz);
  double prev = or_nlres(f);
  f->nl_res0 = prev;
  f->nl_res = prev;
  int32_t sweep;
  for (sweep = 1; sweep <= max_sweeps; sweep++) {
    for (int32_t ucol = 0; ucol < n; ucol++)
      for (int32_t idx = f->Up[ucol]; idx < f->Up[ucol + 1]; idx++) {
        const int32_t urow = f->Ui[idx];
        if (f->uf == 1) {
          float* Ux = (float*)f->Ux;
          float* Lx = (float*)f->Lx;
          float* Di = (float*)f->UDinv;
          const float ax = (float)f->ax[idx];
          Ux[idx] = ax - or_ldotu_f32(f, urow, ucol, urow);
          if (urow == ucol) {
            Di[urow] = 1.0f / Ux[idx];
            Lx[idx] = 1.0f;
          } else {
            Lx[idx] = (ax - or_ldotu_f32(f, ucol, urow, urow)) * Di[urow];
          }
        } else {
          double* Ux = (double*)f->Ux;
          double* Lx = (double*)f->Lx;
          double* Di = (double*)f->UDinv;
          const double ax = f->ax[idx];
          Ux[idx] = ax - or_ldotu_f64(f, urow, ucol, urow);
          if (urow == ucol) {
            Di[urow] = 1.0 / Ux[idx];
            Lx[idx] = 1.0;
          } else {
            Lx[idx] = (ax - or_ldotu_f64(f, ucol, urow, urow)) * Di[urow];
          }
        }
      }
    const double res = or_nlres(f);
    f->nl_res = res;
    if (fabs(res - prev) < tol) break;
    prev = res;
  }
  f->sweeps = sweep > max_sweeps ? max_sweeps : sweep;
}

/* TriangularSolve<double> (:743-774): L y = b by rows (entries left of the
 * diagonal), then U x = y by columns, last to first, with UDinv */
static void or_ilu_solve(const or_ilu* f, double* x) {
  const int32_t n = f->n;
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  for (int32_t row = 0; row < n; row++) {
    double sum = 0.0;
    for (int32_t idx = f->Up[row]; idx < f->Up[row + 1]; idx++) {
      const int32_t col = f->Ui[idx];
      if (col < row) sum += (f->uf == 1 ? (double)((const float*)f->Lx)[idx] : ((const double*)f->Lx)[idx]) * y[col];
    }
    y[row] = x[row] - sum;
  }
  for (int32_t col = n; col-- > 0;) {
    y[col] *= f->uf == 1 ? (double)((const float*)f->UDinv)[col] : ((const double*)f->UDinv)[col];
    for (int32_t idx = f->Up[col]; idx < f->Up[col + 1]; idx++) {
      const int32_t row = f->Ui[idx];
      if (row < col) y[row] -= (f->uf == 1 ? (double)((const float*)f->Ux)[idx] : ((const double*)f->Ux)[idx]) * y[col];
    }
  }
  memcpy(x, y, sizeof(double) * (size_t)n);
  free(y);
}

/* The FGPILU factors of the scaled matrix (values Axs, natural order). */
static or_ilu* or_ilu_build(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Axs, int32_t uf,
                            int32_t ilu_k, int32_t max_sweeps, double tol) {
  or_ilu* f = (or_ilu*)calloc(1, sizeof(or_ilu));
  f->n = n;
  f->uf = uf;
  or_ilu_pattern(n, Ap, Ai, ilu_k, &f->Up, &f->Ui);
  f->nnz = f->Up[n];
  f->ax = (double*)calloc((size_t)(f->nnz > 0 ? f->nnz : 1), sizeof(double));
  /* ValueAt(row, col) (:826-836): A's column col searched for row */
  for (int32_t col = 0; col < n; col++)
    for (int32_t idx = f->Up[col]; idx < f->Up[col + 1]; idx++)
      for (int32_t q = Ap[col]; q < Ap[col + 1]; q++)
        if (Ai[q] == f->Ui[idx]) {
          f->ax[idx] = Axs[q];
          break;
        }
  const size_t fs = uf == 1 ? sizeof(float) : sizeof(double);
  f->Ux = calloc((size_t)(f->nnz > 0 ? f->nnz : 1), fs);
  f->Lx = calloc((size_t)(f->nnz > 0 ? f->nnz : 1), fs);
  f->UDinv = calloc((size_t)n, fs);
  or_fgpilu_run(f, max_sweeps, tol);
  return f;
}

static void or_ilu_free(or_ilu* f) {
  if (!f) return;
  free(f->Up);
  free(f->Ui);
  free(f->ax);
  free(f->Ux);
  free(f->Lx);
  free(f->UDinv);
  free(f);
}

/* Exported: the ILU_k pattern size, and the FGPILU factors of a matrix as
 * given (no scaling) for tests: Up/Ui (n+1 / nnz), Ux/Lx (nnz, double),
 * UDinv (n), info = {sweeps, initial and final nonlinear residual}. Call
 * with Up = NULL for the pattern size only. */
int64_t or_fgpilu(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, int32_t uf, int32_t ilu_k,
                  int32_t max_sweeps, double tol, int32_t* Up, int32_t* Ui, double* Ux, double* Lx, double* UDinv,
                  double* info) {
  or_ilu* f = or_ilu_build(n, Ap, Ai, Ax, uf, ilu_k, Up ? max_sweeps : 0, tol);
  const int64_t nnz = f->nnz;
  if (Up) {
    memcpy(Up, f->Up, sizeof(int32_t) * ((size_t)n + 1));
    memcpy(Ui, f->Ui, sizeof(int32_t) * (size_t)nnz);
    for (int64_t q = 0; q < nnz; q++) {
      Ux[q] = uf == 1 ? (double)((float*)f->Ux)[q] : ((double*)f->Ux)[q];
      Lx[q] = uf == 1 ? (double)((float*)f->Lx)[q] : ((double*)f->Lx)[q];
    }
    for (int32_t i = 0; i < n; i++) UDinv[i] = uf == 1 ? (double)((float*)f->UDinv)[i] : ((double*)f->UDinv)[i];
    if (info) {
      info[0] = f->sweeps;
      info[1] = f->nl_res0;
      info[2] = f->nl_res;
    }
  }
  or_ilu_free(f);
  return nnz;
}

/* x := M^{-1} x with FGPILU factors given in the layout above (double) */
void or_fgpilu_solve(int32_t n, const int32_t* Up, const int32_t* Ui, const double* Ux, const double* Lx,
                     const double* UDinv, double* x) {
  or_ilu f;
  memset(&f, 0, sizeof f);
  f.n = n;
  f.uf = 0;
  f.Up = (int32_t*)Up;
  f.Ui = (int32_t*)Ui;
  f.Ux = (void*)Ux;
  f.Lx = (void*)Lx;
  f.UDinv = (void*)UDinv;
  or_ilu_solve(&f, x);
}

/* GmresLDLIR<UF, double, double> as shipped (:323-634): the reference's
 * diagonal scaling, FGPILU(ILU_k, max_sweeps) of the scaled matrix in its
 * natural order, then Solve with M = the ILU factors. info as or_gmres_ir,
 * plus info[8] = sweeps, info[9] = final nonlinear residual. */
int32_t or_gmres_ilu_ir(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, int32_t uf, int32_t ilu_k,
                        int32_t max_sweeps, int64_t ir_iter, int64_t gmres_iter, double tol, const double* b,
                        double* x, double* info) {
  const double t0 = or_now();
  const int32_t nnz = Ap[n];
  double* Axs = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  double* AD = (double*)malloc(sizeof(double) * (size_t)n);
  or_scaling(n, Ap, Ai, Ax, uf == 0 ? 0 : 1, AD, Axs);
  or_ilu* f = or_ilu_build(n, Ap, Ai, Axs, uf == 0 ? 0 : 1, ilu_k, max_sweeps, tol);
  const double t1 = or_now();
  or_ctx c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  c.Ap = Ap;
  c.Ai = Ai;
  c.Ax = Axs;
  c.ilu = f;
  c.gmres_iter = gmres_iter;
  c.tol = tol;
  double* bs = (double*)malloc(sizeof(double) * (size_t)n);
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* t = (double*)malloc(sizeof(double) * (size_t)n);
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  for (int32_t i = 0; i < n; i++) bs[i] = b[i] * AD[i];
  or_precond_gmres(&c, bs, x);
  or_matmul_upper(n, Ap, Ai, Axs, x, t);
  for (int32_t i = 0; i < n; i++) r[i] = bs[i] - t[i];
  const double bnorm = or_infnrm(n, bs);
  double r_prev = DBL_MAX, r_norm = 0.0;
  int64_t iter;
  int32_t stop = 0;
  for (iter = 1; iter <= ir_iter; iter++) {
    const double scale = or_infnrm(n, r);
    if (scale == 0.0) {
      stop = 3;
      r_norm = 0.0;
      break;
    }
    const double inv = 1 / scale;
    for (int32_t i = 0; i < n; i++) t[i] = r[i] * inv;
    or_precond_gmres(&c, t, d);
    for (int32_t i = 0; i < n; i++) d[i] = d[i] * scale;
    const double d_norm = or_infnrm(n, d), x_norm = or_infnrm(n, x);
    for (int32_t i = 0; i < n; i++) x[i] = x[i] + d[i];
    or_matmul_upper(n, Ap, Ai, Axs, x, t);
    for (int32_t i = 0; i < n; i++) r[i] = bs[i] - t[i];
    r_norm = or_infnrm(n, r) / bnorm;
    if (d_norm <= x_norm * (10 * DBL_EPSILON)) {
      stop = 1;
      break;
    }
    if (fabs(r_norm - r_prev) < tol) {
      stop = 2;
      break;
    }
    r_prev = r_norm;
  }
  for (int32_t i = 0; i < n; i++) x[i] = x[i] * AD[i];
  const double t2 = or_now();
  if (info) {
    info[0] = (double)(iter > ir_iter ? ir_iter : iter);
    info[1] = (double)c.gmres_total;
    info[2] = stop;
    info[3] = 0;
    info[4] = (double)c.precond_count;
    info[5] = r_norm;
    info[6] = t1 - t0;
    info[7] = t2 - t1;
    info[8] = f->sweeps;
    info[9] = f->nl_res;
  }
  or_ilu_free(f);
  free(Axs); free(AD); free(bs); free(r); free(t); free(d);
  return 0;
}
