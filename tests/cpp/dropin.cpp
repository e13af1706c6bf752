// A reference-style caller of QDLDL_etree / QDLDL_factor / QDLDL_solve,
// LUIR_QDLDL and GmresLDLIR<UF,UW,UR>, compiled against include/mpir/*.hpp:
// the calls are
// written exactly as a user of yex33/MPIR writes them (qdldl.hpp,
// gmres_ir.hpp, luir.hpp:79-178 for the QDLDL call sequence).
//
//   dropin_test          full run on a gfx950 device (exit 0 on success)
//   dropin_test --host   no device: etree must work, the GPU entry points
//                        must fail with MPIR_ERR_NODEVICE (exit 0 if so)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mpir/gmres_ir.hpp"
#include "mpir/luir.hpp"
#include "mpir/qdldl.hpp"

#if __cplusplus >= 202002L
// the reference's compile-time checks of Refinable (fp_concepts.hpp:116-127)
static_assert(Refinable<double, double, double>);
static_assert(Refinable<float, double, double>);
static_assert(Refinable<float, float, double>);
static_assert(!Refinable<double, double, float>);
static_assert(!Refinable<double, float, double>);
static_assert(!Refinable<double, float, float>);
static_assert(!Refinable<int, double, double>);
#endif

namespace {

struct Csc {
  int n = 0;
  std::vector<int> p, i;
  std::vector<double> x;
};

// small quasi-definite K = [H A^T; A -D]: H = 2D Laplacian + shift, A random
Csc make_kkt(int k, int m) {
  std::vector<double> h;
  const int n1 = k * k, n = n1 + m;
  std::vector<std::vector<std::pair<int, double>>> cols(n);
  for (int c = 0; c < n1; ++c) {
    const int gx = c % k, gy = c / k;
    if (gy > 0) cols[c].push_back({c - k, -1.0});
    if (gx > 0) cols[c].push_back({c - 1, -1.0});
    cols[c].push_back({c, 4.01});
  }
  unsigned s = 12345u;
  auto rnd = [&]() { s = s * 1664525u + 1013904223u; return (s >> 8) / double(1u << 24); };
  for (int r = 0; r < m; ++r) {
    const int c = n1 + r;
    const int a = static_cast<int>(rnd() * n1), b = (a + 1 + static_cast<int>(rnd() * (n1 - 1))) % n1;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    cols[c].push_back({lo, rnd() - 0.5});
    cols[c].push_back({hi, rnd() - 0.5});
    cols[c].push_back({c, -1e-2});
  }
  Csc K;
  K.n = n;
  K.p.push_back(0);
  for (int c = 0; c < n; ++c) {
    for (auto& e : cols[c]) {
      K.i.push_back(e.first);
      K.x.push_back(e.second);
    }
    K.p.push_back(static_cast<int>(K.i.size()));
  }
  return K;
}

std::vector<double> matvec(const Csc& K, const std::vector<double>& v) {
  std::vector<double> y(K.n, 0.0);
  for (int c = 0; c < K.n; ++c)
    for (int q = K.p[c]; q < K.p[c + 1]; ++q) {
      const int r = K.i[q];
      y[c] += K.x[q] * v[r];
      if (r != c) y[r] += K.x[q] * v[c];
    }
  return y;
}

double rel_residual(const Csc& K, const std::vector<double>& x, const std::vector<double>& b) {
  const std::vector<double> y = matvec(K, x);
  double r = 0, bn = 0;
  for (int i = 0; i < K.n; ++i) {
    r = std::fmax(r, std::fabs(b[i] - y[i]));
    bn = std::fmax(bn, std::fabs(b[i]));
  }
  return r / bn;
}

int failures = 0;
void expect(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++failures;
}

template <typename INT, typename PREC, typename XPREC>
void qdldl_sequence(const Csc& K, double tol, const char* label) {
  // the call sequence of luir.hpp:79-178
  const INT n = K.n;
  std::vector<INT> Ap(K.p.begin(), K.p.end()), Ai(K.i.begin(), K.i.end());
  std::vector<PREC> Ax(K.x.begin(), K.x.end());
  std::vector<INT> work(n), Lnz(n), etree(n);
  const INT sumLnz = QDLDL_etree(n, Ap.data(), Ai.data(), work.data(), Lnz.data(), etree.data());
  expect(sumLnz > 0, "QDLDL_etree returns sum(Lnz)");
  std::vector<INT> Lp(n + 1), Li(sumLnz), iwork(3 * n);
  std::vector<PREC> Lx(sumLnz), D(n), Dinv(n), fwork(n);
  std::vector<unsigned char> bw(n);
  const INT pos = QDLDL_factor(n, Ap.data(), Ai.data(), Ax.data(), Lp.data(), Li.data(), Lx.data(), D.data(),
                               Dinv.data(), Lnz.data(), etree.data(), reinterpret_cast<bool*>(bw.data()),
                               iwork.data(), fwork.data());
  char msg[160];
  std::snprintf(msg, sizeof msg, "%s: QDLDL_factor inertia %lld == %d", label, static_cast<long long>(pos), K.n - 40);
  expect(pos == K.n - 40, msg);
  std::vector<double> xt(K.n);
  for (int i = 0; i < K.n; ++i) xt[i] = std::sin(0.37 * i);
  const std::vector<double> b = matvec(K, xt);
  std::vector<XPREC> x(b.begin(), b.end());
  QDLDL_solve(n, Lp.data(), Li.data(), Lx.data(), Dinv.data(), x.data());
  const double res = rel_residual(K, std::vector<double>(x.begin(), x.end()), b);
  std::snprintf(msg, sizeof msg, "%s: QDLDL_solve relative residual %.2e < %.0e", label, res, tol);
  expect(res < tol, msg);
}

}  // namespace

int main(int argc, char** argv) {
  const bool host_only = argc > 1 && std::strcmp(argv[1], "--host") == 0;
  const Csc K = make_kkt(12, 40);
  if (host_only) {
    std::vector<int> work(K.n), Lnz(K.n), etree(K.n);
    const int s = QDLDL_etree(K.n, K.p.data(), K.i.data(), work.data(), Lnz.data(), etree.data());
    expect(s > 0, "QDLDL_etree runs on the host");
    std::vector<int> Lp(K.n + 1), Li(s);
    std::vector<double> Lx(s), D(K.n), Dinv(K.n);
    const int rc = QDLDL_factor(K.n, K.p.data(), K.i.data(), K.x.data(), Lp.data(), Li.data(), Lx.data(), D.data(),
                                Dinv.data(), Lnz.data(), etree.data(), static_cast<bool*>(nullptr),
                                static_cast<int*>(nullptr), static_cast<double*>(nullptr));
    expect(rc == MPIR_ERR_NODEVICE, "QDLDL_factor without a gfx950 device fails with MPIR_ERR_NODEVICE");
    bool threw = false;
    try {
      GmresLDLIR<float, double, double> solver;
    } catch (const std::runtime_error&) {
      threw = true;
    }
    expect(threw, "GmresLDLIR without a gfx950 device throws");
    return failures == 0 ? 0 : 1;
  }
  qdldl_sequence<int, double, double>(K, 1e-12, "int/double/double");
  qdldl_sequence<long, double, double>(K, 1e-12, "long/double/double");
  qdldl_sequence<int, float, double>(K, 1e-5, "int/float/double");
  qdldl_sequence<long long, float, float>(K, 1e-4, "long long/float/float");

  {
    // LUIR_QDLDL<FACTPREC, RESPREC, WORKPREC, XPREC, INT> with its Stats
    std::vector<double> xt(K.n), x(K.n);
    for (int i = 0; i < K.n; ++i) xt[i] = std::sin(0.11 * i);
    const std::vector<double> b = matvec(K, xt);
    Stats stats;
    LUIR_QDLDL<float, double, double, double, int>(K.n, K.p.data(), K.i.data(), K.x.data(), b.data(), x.data(),
                                                   &stats, 10);
    char msg[160];
    std::snprintf(msg, sizeof msg, "LUIR_QDLDL<float,...>: %d steps, relative residual %.2e < 1e-14",
                  stats.num_refinement_steps_, stats.relative_residual_);
    expect(stats.relative_residual_ < 1e-14 && rel_residual(K, x, b) < 1e-14 && stats.num_refinement_steps_ >= 1,
           msg);
    expect(stats.n_ == K.n && stats.nnz_ == static_cast<long>(K.i.size()) && stats.get_total_time() > 0,
           "LUIR_QDLDL fills Stats");
    std::vector<long> lp(K.p.begin(), K.p.end()), li(K.i.begin(), K.i.end());
    std::vector<double> x2(K.n);
    Stats s2;
    LUIR_QDLDL<double, double, double, double, long>(K.n, lp.data(), li.data(), K.x.data(), b.data(), x2.data(),
                                                     &s2, 0);
    expect(s2.num_refinement_steps_ == 0 && rel_residual(K, x2, b) < 1e-12, "LUIR_QDLDL<double,...> maxiter 0");
  }

  std::vector<std::size_t> Ap(K.p.begin(), K.p.end()), Ai(K.i.begin(), K.i.end());
  std::vector<double> xt(K.n);
  for (int i = 0; i < K.n; ++i) xt[i] = std::cos(0.21 * i);
  const std::vector<double> b = matvec(K, xt);
  GmresLDLIR<float, double, double> solver;
  solver.SetTolerance(1e-12);
  solver.SetMaxIRIterations(10);
  solver.SetMaxGmresIterations(10);
  solver.Compute<double>(Ap, Ai, K.x, 0);
  const std::vector<double> x = solver.Solve(b);
  const double res = rel_residual(K, x, b);
  char msg[160];
  std::snprintf(msg, sizeof msg, "GmresLDLIR<float,double,double>: relative residual %.2e < 1e-13", res);
  expect(res < 1e-13, msg);
  expect(solver.Solve(std::vector<double>(3)).empty(), "Solve on a size mismatch returns an empty vector");
  GmresLDLIR<double, double, double> solver64;
  solver64.Compute<double>(Ap, Ai, K.x, 0);
  const double res64 = rel_residual(K, solver64.Solve(b), b);
  std::snprintf(msg, sizeof msg, "GmresLDLIR<double,double,double>: relative residual %.2e < 1e-13", res64);
  expect(res64 < 1e-13, msg);
  // DumpLog: the VERBOSE log of the last Solve as JSON text
  const std::string log = solver64.DumpLog();
  expect(log.find("\"events\"") != std::string::npos && log.find("\"gmres_ir\"") != std::string::npos &&
             log.find("\"stop_reason\"") != std::string::npos,
         "DumpLog returns the solver log");
  // Refinable<UF, UW, UR> combinations with a float working (and residual)
  // precision: x is held in UW, so it is as accurate as a float can hold
  {
    std::vector<float> bf(b.begin(), b.end());
    GmresLDLIR<float, float, double> sff;
    sff.Compute<double>(Ap, Ai, K.x, 0);
    const std::vector<float> xf = sff.Solve(bf);
    std::vector<double> xfd(xf.begin(), xf.end());
    const double resf = rel_residual(K, xfd, b);
    const mpir_gmres_stats stf = sff.Stats();
    std::snprintf(msg, sizeof msg, "GmresLDLIR<float,float,double>: relative residual %.2e < 1e-5, stop reason %d",
                  resf, stf.stop_reason);
    expect(xf.size() == static_cast<std::size_t>(K.n) && resf < 1e-5 && stf.stop_reason >= 1, msg);
    GmresLDLIR<float, float, float> s3;
    s3.SetTolerance(1e-6f);
    s3.Compute<double>(Ap, Ai, K.x, 0);
    std::vector<double> x3d;
    for (float v : s3.Solve(bf)) x3d.push_back(v);
    const double res3 = rel_residual(K, x3d, b);
    std::snprintf(msg, sizeof msg, "GmresLDLIR<float,float,float>: relative residual %.2e < 1e-5", res3);
    expect(res3 < 1e-5, msg);
    GmresLDLIR<double, double, double> sdd;
    expect(sdd.Solve(b).empty(), "Solve before Compute (n = 0) returns an empty vector");
  }
  // the reference's shipped path: FGPILU with ILU_K = 0 and 5 sweeps, as
  // examples/example.cpp:144-147 calls Compute
  GmresLDLIR<float, double, double> ilu;
  ilu.SetPreconditioner(MPIR_PREC_FGPILU);
  ilu.SetTolerance(1e-12);
  ilu.SetMaxIRIterations(50);
  ilu.Compute<double, 1 << 21, 5, 2048>(Ap, Ai, K.x, 0);
  const double resi = rel_residual(K, ilu.Solve(b), b);
  const mpir_gmres_stats sti = ilu.Stats();
  std::snprintf(msg, sizeof msg, "GmresLDLIR FGPILU ILU(0): %d sweeps, %d IR iterations, relative residual %.2e < 1e-10",
                sti.ilu_sweeps, sti.ir_iterations, resi);
  expect(sti.preconditioner == MPIR_PREC_FGPILU && sti.ilu_sweeps >= 1 && resi < 1e-10, msg);
  return failures == 0 ? 0 : 1;
}
