// Drop-in replacement for GmresLDLIR<UF,UW,UR> of yex33/MPIR
// (include/MPIR/gmres_ir.hpp:90-315) in its LDL^T form
// (docs/Design/SoftDetailedDes/MIS.tex:505-568): Compute() orders the
// matrix, builds the elimination tree and factors P K P^T in UF; Solve()
// runs GMRES-IR with L D L^T as the preconditioner and the residual in UR.
// Same class name, template parameters, member functions and defaults
// (10 IR iterations, 10 GMRES iterations, tolerance 1e-10). The work runs
// on an MI355X (gfx950) through libmpir_hip.so; link with -lmpir_hip.
//
// Precisions: any UF, UW, UR from {float, double} that satisfy the
// reference's Refinable<UF, UW, UR> (fp_concepts.hpp:112-113: floating
// point, eps(UF) >= eps(UW) >= eps(UR)), e.g. <float, float, double>.
// What the reference stores in UW (the matrix values, the scaling vector,
// the scaled right-hand side, the iterate x and the correction d) and in UR
// (the residual) is rounded to that precision on the device, and the
// small-update test uses eps(UW); the Krylov basis and its dot products stay
// in double, at least as accurate as the reference's UW arithmetic.
// Compute() keeps the reference's signature. With the default L D L^T
// preconditioner ILU_k and the template sizes are accepted and ignored;
// after SetPreconditioner(MPIR_PREC_FGPILU) Compute runs the reference's
// shipped FGPILU path (gmres_ir.hpp:323-531) with ILU_k and MAX_SWEEPS.
// Failures throw std::runtime_error with the library's message.
#pragma once

#include <cstddef>
#include <algorithm>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../mpir_hip.h"

namespace mpir_detail {
template <typename T>
constexpr bool is_fp_v = std::is_same<T, float>::value || std::is_same<T, double>::value;
// fp_concepts.hpp:112-113 for the precisions the device supports
template <typename UF, typename UW, typename UR>
constexpr bool refinable_v = is_fp_v<UF> && is_fp_v<UW> && is_fp_v<UR> &&
                             std::numeric_limits<UF>::epsilon() >= std::numeric_limits<UW>::epsilon() &&
                             std::numeric_limits<UW>::epsilon() >= std::numeric_limits<UR>::epsilon();
template <typename T>
constexpr int tag_v = std::is_same<T, float>::value ? MPIR_F32 : MPIR_F64;
}  // namespace mpir_detail

#if __cplusplus >= 202002L
template <typename UF, typename UW, typename UR>
concept Refinable = mpir_detail::refinable_v<UF, UW, UR>;
#define MPIR_REQUIRES_REFINABLE requires Refinable<UF, UW, UR>
#else
#define MPIR_REQUIRES_REFINABLE
#endif

template <typename UF, typename UW, typename UR>
MPIR_REQUIRES_REFINABLE class GmresLDLIR {
  static_assert(mpir_detail::refinable_v<UF, UW, UR>,
                "GmresLDLIR<UF, UW, UR>: float or double with eps(UF) >= eps(UW) >= eps(UR) (Refinable)");

 public:
  GmresLDLIR() {
    check(mpir_gmres_ldl_create(mpir_detail::tag_v<UF>, mpir_detail::tag_v<UW>, mpir_detail::tag_v<UR>, &h_),
          "GmresLDLIR");
  }
  ~GmresLDLIR() { mpir_gmres_ldl_destroy(h_); }
  GmresLDLIR(const GmresLDLIR&) = delete;
  GmresLDLIR& operator=(const GmresLDLIR&) = delete;

  // Upper-triangular CSC of K (rows ascending, diagonal present in every
  // column). Repeated calls with the same pattern reuse the symbolic phase.
  template <typename TAx, std::size_t MAX_MATRIX_SIZE = 1 << 21, std::size_t MAX_SWEEPS = 10,
            std::size_t BATCH_SIZE = 4096>
  void Compute(std::vector<std::size_t> Ap, std::vector<std::size_t> Ai, std::vector<TAx> Ax, std::size_t ILU_k = 0) {
    if (prec_ == MPIR_PREC_FGPILU && (ILU_k != ilu_k_ || MAX_SWEEPS != sweeps_)) {
      check(mpir_gmres_ldl_set_preconditioner(h_, prec_, static_cast<int32_t>(ILU_k), static_cast<int32_t>(MAX_SWEEPS)),
            "Compute");
      ilu_k_ = ILU_k;
      sweeps_ = MAX_SWEEPS;
    }
    if (Ap.empty() || Ap.size() - 1 > static_cast<std::size_t>(std::numeric_limits<int32_t>::max()) ||
        Ai.size() > static_cast<std::size_t>(std::numeric_limits<int32_t>::max()) || Ax.size() != Ai.size())
      throw std::invalid_argument("GmresLDLIR::Compute: bad CSC sizes");
    n_ = Ap.size() - 1;
    std::vector<int32_t> ap(Ap.begin(), Ap.end()), ai(Ai.begin(), Ai.end());
    std::vector<double> ax(Ax.begin(), Ax.end());
    check(mpir_gmres_ldl_compute(h_, static_cast<int32_t>(n_), ap.data(), ai.data(), ax.data()), "Compute");
  }

  std::vector<UW> Solve(const std::vector<UW>& b) {
    if (b.size() != n_) return {};  // as PrecondGmres on a size mismatch
    std::vector<double> bd(b.begin(), b.end()), xd(n_);
    check(mpir_gmres_ldl_solve(h_, bd.data(), xd.data(), 1, static_cast<int64_t>(n_)), "Solve");
    return std::vector<UW>(xd.begin(), xd.end());
  }

  void SetMaxIRIterations(std::size_t n) { check(mpir_gmres_ldl_set_max_ir(h_, static_cast<int64_t>(n)), "SetMaxIRIterations"); }
  void SetMaxGmresIterations(std::size_t n) {
    check(mpir_gmres_ldl_set_max_gmres(h_, static_cast<int64_t>(n)), "SetMaxGmresIterations");
  }
  void SetTolerance(UR tol) { check(mpir_gmres_ldl_set_tol(h_, static_cast<double>(tol)), "SetTolerance"); }

  // The solver log of the reference's VERBOSE build (gmres_ir.hpp:305-314)
  // as JSON text: max_outer_iterations, max_inner_iterations, tolerance and
  // the "gmres_ir" event of the last Solve. The reference returns an
  // nlohmann::json; parse this string with any JSON library to get one.
  std::string DumpLog() const {
    const int64_t len = mpir_gmres_ldl_log(h_, nullptr, 0);
    check(static_cast<int>(len < 0 ? len : 0), "DumpLog");
    std::string out(static_cast<std::size_t>(len) + 1, '\0');
    check(static_cast<int>(std::min<int64_t>(0, mpir_gmres_ldl_log(h_, &out[0], len + 1))), "DumpLog");
    out.resize(static_cast<std::size_t>(len));
    return out;
  }

  // Additions of this implementation (not in the reference's class).
  void SetOrdering(int ordering) { check(mpir_gmres_ldl_set_ordering(h_, ordering), "SetOrdering"); }
  // MPIR_PREC_LDL (default) or MPIR_PREC_FGPILU; takes effect at the next Compute
  void SetPreconditioner(int kind) {
    check(mpir_gmres_ldl_set_preconditioner(h_, kind, 0, 10), "SetPreconditioner");
    prec_ = kind;
    ilu_k_ = 0;
    sweeps_ = 10;
  }
  mpir_gmres_stats Stats() const {
    mpir_gmres_stats st{};
    check(mpir_gmres_ldl_stats(h_, &st), "Stats");
    return st;
  }

 private:
  static void check(int rc, const char* what) {
    if (rc < 0)
      throw std::runtime_error(std::string("GmresLDLIR::") + what + " failed (" + std::to_string(rc) +
                               "): " + mpir_last_error());
  }
  mpir_gmres_ldl_t h_ = nullptr;
  std::size_t n_ = 0;
  int prec_ = MPIR_PREC_LDL;
  std::size_t ilu_k_ = 0, sweeps_ = 10;
};
