// mtx_solve: the reference's example driver (examples/example.cpp:65-196 of
// yex33/MPIR) on the MI355X path.
//
//   mtx_solve <matrix.mtx> [--uf f64|f32] [--ordering natural|amd|nd] [--log]
//
// Reads a symmetric Matrix Market file (lower triangle, read transposed into
// the upper CSC), builds GmresLDLIR<UF, double, double> with tolerance 1e-15,
// 200 IR and 50 GMRES iterations, factors, solves K x = K 1 and reports
// ||x - 1|| / ||1|| and ||b - K x|| / ||b|| (infinity norms), with the
// per-iteration relative residuals the reference's Solve prints. --log
// writes <stem>_<timestamp>.json beside the matrix in the schema of the
// reference's VERBOSE log. Differences from the reference: the residual
// precision is double (the reference's example uses dd_real from the QD
// library), and a fill-reducing ordering is applied (default AMD; natural
// keeps the matrix as given, as the reference does).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "mpir_hip.h"

namespace {

[[noreturn]] void die(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, mpir_last_error());
  std::exit(EXIT_FAILURE);
}

double inf_norm(const std::vector<double>& v) {
  double m = 0.0;
  for (double e : v) m = std::fmax(m, std::fabs(e));
  return m;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("Usage: %s <matrix_file_path> [--uf f64|f32] [--ordering natural|amd|nd] [--log]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const std::string path = argv[1];
  int uf = MPIR_F64, ordering = MPIR_ORDER_AMD;
  bool want_log = false;
  for (int a = 2; a < argc; ++a) {
    const std::string o = argv[a];
    if (o == "--uf" && a + 1 < argc) {
      const std::string v = argv[++a];
      uf = v == "f32" ? MPIR_F32 : MPIR_F64;
    } else if (o == "--ordering" && a + 1 < argc) {
      const std::string v = argv[++a];
      ordering = v == "natural" ? MPIR_ORDER_NATURAL : v == "nd" ? MPIR_ORDER_ND : MPIR_ORDER_AMD;
    } else if (o == "--log") {
      want_log = true;
    } else {
      std::printf("unknown option %s\n", o.c_str());
      return EXIT_FAILURE;
    }
  }
  int32_t n = 0, sym = 0;
  int64_t nnz = 0;
  if (mpir_read_mtx_upper(path.c_str(), &n, &nnz, &sym, nullptr, nullptr, nullptr) < 0) {
    std::printf("Error: cannot read the matrix file at %s (%s)\n", path.c_str(), mpir_last_error());
    return EXIT_FAILURE;
  }
  std::vector<int32_t> Ap(n + 1), Ai(nnz);
  std::vector<double> Ax(nnz);
  if (mpir_read_mtx_upper(path.c_str(), &n, &nnz, &sym, Ap.data(), Ai.data(), Ax.data()) < 0) die("read");
  std::printf("%s\n", sym ? "symmetric" : "non symmetric");

  mpir_gmres_ldl_t h = nullptr;
  if (mpir_gmres_ldl_create(uf, MPIR_F64, MPIR_F64, &h) < 0) die("GmresLDLIR");
  if (mpir_gmres_ldl_set_tol(h, 1e-15) < 0 || mpir_gmres_ldl_set_max_ir(h, 200) < 0 ||
      mpir_gmres_ldl_set_max_gmres(h, 50) < 0 || mpir_gmres_ldl_set_ordering(h, ordering) < 0)
    die("configure");
  auto t0 = std::chrono::high_resolution_clock::now();
  if (mpir_gmres_ldl_compute(h, n, Ap.data(), Ai.data(), Ax.data()) < 0) die("Compute");
  auto t1 = std::chrono::high_resolution_clock::now();
  std::printf("%-30s: %lld ms\n", "factorization time",
              static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()));

  const std::vector<double> x_ref(n, 1.0);
  std::vector<double> b(n), x(n), Kx(n);
  if (mpir_spmv_upper_f64(n, Ap.data(), Ai.data(), Ax.data(), x_ref.data(), b.data()) < 0) die("MatrixMultiply");
  t0 = std::chrono::high_resolution_clock::now();
  if (mpir_gmres_ldl_solve(h, b.data(), x.data(), 1, n) < 0) die("Solve");
  t1 = std::chrono::high_resolution_clock::now();

  // the relative residual of each refinement step, as the reference's Solve prints
  const int64_t len = mpir_gmres_ldl_log(h, nullptr, 0);
  std::string log(static_cast<size_t>(len > 0 ? len : 0) + 1, '\0');
  if (len > 0) mpir_gmres_ldl_log(h, &log[0], len + 1);
  log.resize(static_cast<size_t>(len > 0 ? len : 0));
  for (size_t pos = log.find("{\"iteration\": "); pos != std::string::npos;
       pos = log.find("{\"iteration\": ", pos + 1)) {
    const int it = std::atoi(log.c_str() + pos + 14);
    const size_t rp = log.find("\"residual\": ", pos);
    std::printf("relative residual @ iter %d = %s\n", it,
                log.substr(rp + 12, log.find(',', rp) - rp - 12).c_str());
  }
  std::printf("%-30s: %lld ms\n", "solve time",
              static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()));

  std::vector<double> diff(n);
  for (int32_t i = 0; i < n; ++i) diff[i] = x[i] - x_ref[i];
  const double relative_error = inf_norm(diff) / inf_norm(x_ref);
  if (mpir_spmv_upper_f64(n, Ap.data(), Ai.data(), Ax.data(), x.data(), Kx.data()) < 0) die("MatrixMultiply");
  for (int32_t i = 0; i < n; ++i) diff[i] = b[i] - Kx[i];
  const double relative_residual = inf_norm(diff) / inf_norm(b);
  std::printf("%-30s: %.17g\n", "||x - xref|| / ||xref||", relative_error);
  std::printf("%-30s: %.17g\n", "||b - Ax|| / ||b||", relative_residual);

  if (want_log) {
    // the example's additions to the solver log (examples/example.cpp:175-195)
    std::string stem = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
    if (stem.size() > 4 && stem.compare(stem.size() - 4, 4, ".mtx") == 0) stem.resize(stem.size() - 4);
    char extra[1024];
    std::snprintf(extra, sizeof extra,
                  "  \"matrix\": \"%s\",\n  \"factorization_precision\": \"%s\",\n"
                  "  \"working_precision\": \"double\",\n  \"residual_precision\": \"double\",\n"
                  "  \"relative_error\": %.17g,\n  \"relative_residual\": %.17g,\n  \"ilu_k\": 0,\n",
                  stem.c_str(), uf == MPIR_F32 ? "float" : "double", relative_error, relative_residual);
    const std::string full = "{\n" + std::string(extra) + log.substr(log.find('{') + 2);
    std::time_t now = std::time(nullptr);
    char ts[32];
    std::strftime(ts, sizeof ts, "%Y-%m-%d-%H-%M-%S", std::localtime(&now));
    const std::string dir = path.find_last_of('/') == std::string::npos ? "." : path.substr(0, path.find_last_of('/'));
    const std::string out = dir + "/" + stem + "_" + ts + ".json";
    std::FILE* f = std::fopen(out.c_str(), "w");
    if (!f) {
      std::perror(out.c_str());
      return EXIT_FAILURE;
    }
    std::fputs(full.c_str(), f);
    std::fclose(f);
    std::printf("%-30s: %s\n", "log", out.c_str());
  }
  mpir_gmres_ldl_destroy(h);
  return EXIT_SUCCESS;
}
