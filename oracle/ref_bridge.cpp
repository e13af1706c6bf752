// ORACLE -- test infrastructure only. extern "C" entry points over the
// reference's own headers, compiled as shipped from /root/reference by
// oracle/Makefile (target "ref") into oracle/_ref/libqdldl_ref.so:
//   include/MPIR/qdldl.hpp  QDLDL_etree / QDLDL_factor / QDLDL_solve
//   include/MPIR/ops.hpp    MatrixMultiply / VectorDot / Dnrm2 / InfNrm
// gmres_ir.hpp is not built: it needs <print>, which this toolchain lacks.
#include <cstddef>
#include <cstdint>
#include <vector>

#include "ops.hpp"
#include "qdldl.hpp"

extern "C" {

int32_t ref_etree_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* work, int32_t* Lnz, int32_t* etree) {
  return QDLDL_etree<int32_t>(n, Ap, Ai, work, Lnz, etree);
}

int32_t ref_factor_f64_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, int32_t* Lp, int32_t* Li,
                           double* Lx, double* D, double* Dinv, const int32_t* Lnz, const int32_t* etree) {
  std::vector<unsigned char> b(static_cast<size_t>(n) * sizeof(bool) + 1);
  std::vector<int32_t> iw(3 * static_cast<size_t>(n));
  std::vector<double> fw(static_cast<size_t>(n));
  return QDLDL_factor<int32_t, double>(n, Ap, Ai, Ax, Lp, Li, Lx, D, Dinv, Lnz, etree,
                                       reinterpret_cast<bool*>(b.data()), iw.data(), fw.data());
}

int32_t ref_factor_f32_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const float* Ax, int32_t* Lp, int32_t* Li,
                           float* Lx, float* D, float* Dinv, const int32_t* Lnz, const int32_t* etree) {
  std::vector<unsigned char> b(static_cast<size_t>(n) * sizeof(bool) + 1);
  std::vector<int32_t> iw(3 * static_cast<size_t>(n));
  std::vector<float> fw(static_cast<size_t>(n));
  return QDLDL_factor<int32_t, float>(n, Ap, Ai, Ax, Lp, Li, Lx, D, Dinv, Lnz, etree,
                                      reinterpret_cast<bool*>(b.data()), iw.data(), fw.data());
}

void ref_solve_f64_f64(int32_t n, const int32_t* Lp, const int32_t* Li, const double* Lx, const double* Dinv,
                       double* x) {
  QDLDL_solve<int32_t, double, double>(n, Lp, Li, Lx, Dinv, x);
}
void ref_solve_f32_f64(int32_t n, const int32_t* Lp, const int32_t* Li, const float* Lx, const float* Dinv, double* x) {
  QDLDL_solve<int32_t, float, double>(n, Lp, Li, Lx, Dinv, x);
}
void ref_solve_f32_f32(int32_t n, const int32_t* Lp, const int32_t* Li, const float* Lx, const float* Dinv, float* x) {
  QDLDL_solve<int32_t, float, float>(n, Lp, Li, Lx, Dinv, x);
}

void ref_matmul_upper(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x, double* y) {
  std::vector<std::size_t> p(Ap, Ap + n + 1), i(Ai, Ai + Ap[n]);
  std::vector<double> v(Ax, Ax + Ap[n]), xv(x, x + n);
  std::vector<double> r = MatrixMultiply<double>(p, i, v, xv);
  for (int32_t k = 0; k < n; ++k) y[k] = r[k];
}

double ref_dot(int64_t n, const double* a, const double* b) {
  std::vector<double> av(a, a + n), bv(b, b + n);
  return VectorDot<double>(av, bv);
}

double ref_dnrm2(int64_t n, const double* x) {
  std::vector<double> v(x, x + n);
  return Dnrm2<double>(v);
}

double ref_infnrm(int64_t n, const double* x) {
  std::vector<double> v(x, x + n);
  return InfNrm<double>(v);
}

}  // extern "C"
