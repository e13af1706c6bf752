// Drop-in replacements for the QDLDL entry points of yex33/MPIR
// (include/MPIR/qdldl.hpp): QDLDL_etree (:48-105), QDLDL_factor (:172-345)
// and QDLDL_solve (:417-432), with the same names, template parameters,
// arguments and return values. The work runs on an MI355X (gfx950) through
// libmpir_hip.so (include/mpir_hip.h); link with -lmpir_hip.
//
// Supported types: INT of 32 or 64 bits (signed); PREC / FACTPREC float or
// double; XPREC double, or float with a float factor.
// Return values follow QDLDL: etree returns sum(Lnz), -1 for an invalid
// column, -2 on overflow; factor returns the number of positive pivots or -1
// on a zero pivot. Other failures (no gfx950 device, HIP errors) return the
// library's negative status; QDLDL_solve, which returns nothing in QDLDL,
// throws std::runtime_error instead.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../mpir_hip.h"

namespace mpir_dropin {

template <class INT>
constexpr bool kIndex = std::is_integral<INT>::value && std::is_signed<INT>::value &&
                        (sizeof(INT) == 4 || sizeof(INT) == 8);
template <class T>
constexpr bool kReal = std::is_same<T, float>::value || std::is_same<T, double>::value;

inline void throw_status(long long rc, const char* what) {
  throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + mpir_last_error());
}

}  // namespace mpir_dropin

template <typename INT>
INT QDLDL_etree(const INT n, const INT* Ap, const INT* Ai, INT* work, INT* Lnz, INT* etree) {
  static_assert(mpir_dropin::kIndex<INT>, "INT must be a signed 32- or 64-bit integer");
  if constexpr (sizeof(INT) == 4) {
    return static_cast<INT>(mpir_qdldl_etree_i32(static_cast<int32_t>(n), reinterpret_cast<const int32_t*>(Ap),
                                                 reinterpret_cast<const int32_t*>(Ai),
                                                 reinterpret_cast<int32_t*>(work), reinterpret_cast<int32_t*>(Lnz),
                                                 reinterpret_cast<int32_t*>(etree)));
  } else {
    return static_cast<INT>(mpir_qdldl_etree_i64(static_cast<int64_t>(n), reinterpret_cast<const int64_t*>(Ap),
                                                 reinterpret_cast<const int64_t*>(Ai),
                                                 reinterpret_cast<int64_t*>(work), reinterpret_cast<int64_t*>(Lnz),
                                                 reinterpret_cast<int64_t*>(etree)));
  }
}

template <typename INT, typename PREC>
INT QDLDL_factor(INT n, const INT* Ap, const INT* Ai, const PREC* Ax, INT* Lp, INT* Li, PREC* Lx, PREC* D,
                 PREC* Dinv, const INT* Lnz, const INT* etree, bool* bwork, INT* iwork, PREC* fwork) {
  static_assert(mpir_dropin::kIndex<INT>, "INT must be a signed 32- or 64-bit integer");
  static_assert(mpir_dropin::kReal<PREC>, "PREC must be float or double");
  if constexpr (sizeof(INT) == 4) {
    using I = int32_t;
    auto ap = reinterpret_cast<const I*>(Ap);
    auto ai = reinterpret_cast<const I*>(Ai);
    auto lp = reinterpret_cast<I*>(Lp);
    auto li = reinterpret_cast<I*>(Li);
    auto lnz = reinterpret_cast<const I*>(Lnz);
    auto et = reinterpret_cast<const I*>(etree);
    auto iw = reinterpret_cast<I*>(iwork);
    if constexpr (std::is_same<PREC, double>::value)
      return static_cast<INT>(mpir_qdldl_factor_f64_i32(static_cast<I>(n), ap, ai, Ax, lp, li, Lx, D, Dinv, lnz, et,
                                                        reinterpret_cast<unsigned char*>(bwork), iw, fwork));
    else
      return static_cast<INT>(mpir_qdldl_factor_f32_i32(static_cast<I>(n), ap, ai, Ax, lp, li, Lx, D, Dinv, lnz, et,
                                                        reinterpret_cast<unsigned char*>(bwork), iw, fwork));
  } else {
    using I = int64_t;
    auto ap = reinterpret_cast<const I*>(Ap);
    auto ai = reinterpret_cast<const I*>(Ai);
    auto lp = reinterpret_cast<I*>(Lp);
    auto li = reinterpret_cast<I*>(Li);
    auto lnz = reinterpret_cast<const I*>(Lnz);
    auto et = reinterpret_cast<const I*>(etree);
    auto iw = reinterpret_cast<I*>(iwork);
    if constexpr (std::is_same<PREC, double>::value)
      return static_cast<INT>(mpir_qdldl_factor_f64_i64(static_cast<I>(n), ap, ai, Ax, lp, li, Lx, D, Dinv, lnz, et,
                                                        reinterpret_cast<unsigned char*>(bwork), iw, fwork));
    else
      return static_cast<INT>(mpir_qdldl_factor_f32_i64(static_cast<I>(n), ap, ai, Ax, lp, li, Lx, D, Dinv, lnz, et,
                                                        reinterpret_cast<unsigned char*>(bwork), iw, fwork));
  }
}

template <typename INT, typename FACTPREC, typename XPREC>
void QDLDL_solve(const INT n, const INT* Lp, const INT* Li, const FACTPREC* Lx, const FACTPREC* Dinv, XPREC* x) {
  static_assert(mpir_dropin::kIndex<INT>, "INT must be a signed 32- or 64-bit integer");
  static_assert(mpir_dropin::kReal<FACTPREC> && mpir_dropin::kReal<XPREC>, "FACTPREC / XPREC must be float or double");
  static_assert(sizeof(XPREC) >= sizeof(FACTPREC), "XPREC must be at least as precise as FACTPREC");
  long long rc;
  if constexpr (sizeof(INT) == 4) {
    using I = int32_t;
    auto lp = reinterpret_cast<const I*>(Lp);
    auto li = reinterpret_cast<const I*>(Li);
    if constexpr (std::is_same<FACTPREC, double>::value)
      rc = mpir_qdldl_solve_f64_f64_i32(static_cast<I>(n), lp, li, Lx, Dinv, x);
    else if constexpr (std::is_same<XPREC, double>::value)
      rc = mpir_qdldl_solve_f32_f64_i32(static_cast<I>(n), lp, li, Lx, Dinv, x);
    else
      rc = mpir_qdldl_solve_f32_f32_i32(static_cast<I>(n), lp, li, Lx, Dinv, x);
  } else {
    using I = int64_t;
    auto lp = reinterpret_cast<const I*>(Lp);
    auto li = reinterpret_cast<const I*>(Li);
    if constexpr (std::is_same<FACTPREC, double>::value)
      rc = mpir_qdldl_solve_f64_f64_i64(static_cast<I>(n), lp, li, Lx, Dinv, x);
    else if constexpr (std::is_same<XPREC, double>::value)
      rc = mpir_qdldl_solve_f32_f64_i64(static_cast<I>(n), lp, li, Lx, Dinv, x);
    else
      rc = mpir_qdldl_solve_f32_f32_i64(static_cast<I>(n), lp, li, Lx, Dinv, x);
  }
  if (rc < 0) mpir_dropin::throw_status(rc, "QDLDL_solve");
}
