// Drop-in replacement for LUIR_QDLDL and Stats of yex33/MPIR
// (include/qdldl-luir/luir.hpp:33-241, include/qdldl-luir/stats.hpp:6-114):
// QDLDL factorization of the matrix as given in FACTPREC, then classic
// iterative refinement with the residual in RESPREC. Same function name,
// template parameters, arguments and Stats fields. Factorization and
// refinement run on an MI355X (gfx950) through libmpir_hip.so; link with
// -lmpir_hip.
//
// Supported precisions: FACTPREC float or double; RESPREC, WORKPREC and
// XPREC double (the quad types of the reference need the QD library). The
// reference prints a message and exits on invalid input; this version
// prints the same message and throws std::runtime_error.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <typeinfo>

#include "../mpir_hip.h"

class Stats {
 public:
  Stats() { init(); }
  void init() {
    n_ = -1;
    nnz_ = -1;
    time_sym_ = -1;
    time_num_fact_ = -1;
    time_solve_ = -1;
    time_ref_ = 0;
    num_refinement_steps_ = -1;
    Lnz_ = static_cast<unsigned long>(-1);
    relative_residual_ = -1;
  }
  double get_total_time() { return time_sym_ + time_num_fact_ + time_solve_ + time_ref_; }
  double get_time_num_fact() { return time_num_fact_; }
  int get_num_refinement_steps() { return num_refinement_steps_; }
  double get_residual() { return relative_residual_; }

  std::string fact_prec_, work_prec_, res_prec_;
  int n_;
  long int nnz_;
  double time_sym_, time_num_fact_, time_solve_, time_ref_;
  int num_refinement_steps_;
  unsigned long Lnz_;
  double relative_residual_;

  void print(const char* str = nullptr) const {
    auto name = [](const std::string& p) {
      return p == typeid(float).name() ? "single" : p == typeid(double).name() ? "double" : "unknown";
    };
    std::printf("------------------------------------------------------------------\n");
    std::printf("%-20s%s%s\n", "Matrix", str ? ": " : "", str ? str : "");
    std::printf("%-20s: %d\n", "   size", n_);
    std::printf("%-20s: %ld\n", "   nnz in triu(A)", nnz_);
    std::printf("%-20s: %lu\n", "   nnz in L", Lnz_);
    std::printf("%-20s: %s,  %s, %s\n", "precisions", name(fact_prec_), name(work_prec_), name(res_prec_));
    std::printf("%-20s: %.2e\n", "||b-A*x||/||b||", relative_residual_);
    std::printf("%-20s: %d\n", "Refinement steps", num_refinement_steps_);
    std::printf("%-20s: %.2e\n", "TOTAL TIME", time_sym_ + time_num_fact_ + time_solve_ + time_ref_);
    std::printf("%-20s: %.2e\n", "  num fact", time_num_fact_);
    std::printf("%-20s: %.2e\n", "  refine", time_ref_);
    std::printf("%-20s: %.2e\n", "  solve", time_solve_);
    std::printf("%-20s: %.2e\n", "  symbolic", time_sym_);
    std::printf("------------------------------------------------------------------\n");
  }
};

template <typename FACTPREC, typename RESPREC, typename WORKPREC, typename XPREC, typename INT>
void LUIR_QDLDL(INT n, const INT* Ap, const INT* Ai, const WORKPREC* Ax, const WORKPREC* b, XPREC* x, Stats* stats,
                int maxiter = 10) {
  static_assert(std::is_same<FACTPREC, float>::value || std::is_same<FACTPREC, double>::value,
                "FACTPREC must be float or double");
  static_assert(std::is_same<RESPREC, double>::value && std::is_same<WORKPREC, double>::value &&
                    std::is_same<XPREC, double>::value,
                "RESPREC, WORKPREC and XPREC must be double");
  static_assert(std::is_integral<INT>::value && std::is_signed<INT>::value && (sizeof(INT) == 4 || sizeof(INT) == 8),
                "INT must be a signed 32- or 64-bit integer");
  mpir_luir_stats st{};
  long long rc;
  if constexpr (sizeof(INT) == 4) {
    auto ap = reinterpret_cast<const int32_t*>(Ap);
    auto ai = reinterpret_cast<const int32_t*>(Ai);
    if constexpr (std::is_same<FACTPREC, double>::value)
      rc = mpir_luir_qdldl_f64_i32(static_cast<int32_t>(n), ap, ai, Ax, b, x, &st, maxiter);
    else
      rc = mpir_luir_qdldl_f32_i32(static_cast<int32_t>(n), ap, ai, Ax, b, x, &st, maxiter);
  } else {
    auto ap = reinterpret_cast<const int64_t*>(Ap);
    auto ai = reinterpret_cast<const int64_t*>(Ai);
    if constexpr (std::is_same<FACTPREC, double>::value)
      rc = mpir_luir_qdldl_f64_i64(static_cast<int64_t>(n), ap, ai, Ax, b, x, &st, maxiter);
    else
      rc = mpir_luir_qdldl_f32_i64(static_cast<int64_t>(n), ap, ai, Ax, b, x, &st, maxiter);
  }
  if (rc < 0) throw std::runtime_error("LUIR_QDLDL failed (" + std::to_string(rc) + "): " + mpir_last_error());
  if (stats) {
    stats->fact_prec_ = typeid(FACTPREC).name();
    stats->work_prec_ = typeid(WORKPREC).name();
    stats->res_prec_ = typeid(RESPREC).name();
    stats->n_ = st.n;
    stats->nnz_ = static_cast<long int>(st.nnz);
    stats->Lnz_ = static_cast<unsigned long>(st.lnz);
    stats->num_refinement_steps_ = st.num_refinement_steps;
    stats->relative_residual_ = st.relative_residual;
    stats->time_sym_ = st.time_sym;
    stats->time_num_fact_ = st.time_num_fact;
    stats->time_solve_ = st.time_solve;
    stats->time_ref_ = st.time_ref;
  }
}
