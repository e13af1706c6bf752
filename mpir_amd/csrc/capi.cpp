// extern "C" entry points of libmpir_hip.so (declared in include/mpir_hip.h).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpir_hip.h"
#include "common.hpp"
#include "host/sparse.hpp"
#include "serve.hpp"
#include "solver.hpp"

namespace mpir {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return kErrInternal;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return kErrInternal;
  }
}

// The QDLDL drop-ins keep one engine per factor precision, reused while the
// pattern stays the same (compared entry by entry). The reference's QDLDL
// functions are reentrant (qdldl.hpp:22-23: all state in the arguments), so
// every calling thread has engines of its own, each with its own stream and
// buffers: calls from different threads run concurrently and never share a
// handle.
struct QdldlEngine {
  std::unique_ptr<GmresLdl> solver;
  std::vector<int32_t> Ap, Ai;  // the pattern the solver was analysed for
  bool same_pattern(int32_t n, const int32_t* p, const int32_t* i) const {
    return solver && solver->analyzed() && Ap.size() == static_cast<size_t>(n) + 1 && std::equal(p, p + n + 1, Ap.begin()) &&
           std::equal(i, i + p[n], Ai.begin());
  }
  void analyze(int32_t n, const int32_t* p, const int32_t* i) {
    Ap.clear();
    Ai.clear();
    solver->analyze(n, p, i);
    Ap.assign(p, p + n + 1);
    Ai.assign(i, i + p[n]);
  }
};
struct QdldlEngines {
  QdldlEngine factor[4], luir[2], solve[2];
};
QdldlEngines& engines() {
  thread_local QdldlEngines e;
  return e;
}

bool narrow(int64_t v, int32_t* out) {
  if (v < std::numeric_limits<int32_t>::min() || v > std::numeric_limits<int32_t>::max()) return false;
  *out = static_cast<int32_t>(v);
  return true;
}

std::vector<int32_t> narrow_vec(const int64_t* p, int64_t n) {
  std::vector<int32_t> v(n);
  for (int64_t i = 0; i < n; ++i)
    if (!narrow(p[i], &v[i])) throw Error(kErrInvalid, "index does not fit in int32");
  return v;
}

int32_t etree_impl(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* work, int32_t* Lnz, int32_t* etree,
                   bool check_i32, int64_t* total) {
  // the reference reports entries below the diagonal on stdout
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q)
      if (Ai[q] > j) {
        bool empty_before = false;
        for (int32_t c = 0; c < n; ++c)
          if (Ap[c] == Ap[c + 1]) { empty_before = true; break; }
        if (!empty_before) {
          std::printf("!!! Matrix is not upper triangular\n");
          std::fflush(stdout);
        }
        j = n;
        break;
      }
  *total = etree_counts(n, Ap, Ai, work, Lnz, etree, check_i32);
  return 0;
}

// A 16-bit value as its IEEE binary16 (P = kF16) or bfloat16 (P = kBF16)
// bit pattern, the storage of the 16-bit QDLDL entry points.
template <int P>
struct Bits16 {
  uint16_t v;
};
static_assert(sizeof(Bits16<kF16>) == 2, "Bits16 is a bare uint16_t");
template <class T>
struct ValueIO {
  static constexpr int prec = std::is_same<T, double>::value ? kF64 : kF32;
  static double get(T x) { return static_cast<double>(x); }
  static T put(double v) { return static_cast<T>(v); }
};
template <>
struct ValueIO<Bits16<kF16>> {
  static constexpr int prec = kF16;
  static double get(Bits16<kF16> x) { return __half2float(__builtin_bit_cast(__half, x.v)); }
  static Bits16<kF16> put(double v) { return {__builtin_bit_cast(uint16_t, __float2half(static_cast<float>(v)))}; }
};
template <>
struct ValueIO<Bits16<kBF16>> {
  static constexpr int prec = kBF16;
  static double get(Bits16<kBF16> x) { return __bfloat162float(__builtin_bit_cast(__hip_bfloat16, x.v)); }
  static Bits16<kBF16> put(double v) {
    return {__builtin_bit_cast(uint16_t, __float2bfloat16(static_cast<float>(v)))};
  }
};

// Factor the given (already ordered) upper CSC on the GPU with QDLDL's
// semantics and write QDLDL's CSC output. 16-bit factors are computed with
// FP32 arithmetic, every matrix value and L entry rounded to the format as
// it is stored; D and Dinv (FP32 inside) are rounded on output.
template <class T>
int64_t factor_impl(int32_t n, const int32_t* Ap, const int32_t* Ai, const T* Ax, int64_t* Lp64, int32_t* Li, T* Lx,
                    T* D, T* Dinv, const int32_t* Lnz, const int32_t* etree) {
  const int prec = ValueIO<T>::prec;
  QdldlEngine& E = engines().factor[prec];
  if (!E.solver) {
    E.solver = std::make_unique<GmresLdl>(prec, kF64, kF64);
    E.solver->set_ordering(kOrderNatural);
    E.solver->set_scaling(kScaleNone);
  }
  const int64_t nnz = Ap[n];
  if (!E.same_pattern(n, Ap, Ai)) E.analyze(n, Ap, Ai);
  // the caller's etree / counts must describe this pattern
  std::vector<int32_t> work(n), lnz(n), par(n);
  etree_counts(n, Ap, Ai, work.data(), lnz.data(), par.data(), false);
  for (int32_t j = 0; j < n; ++j)
    if (lnz[j] != Lnz[j] || par[j] != etree[j]) throw Error(kErrInvalid, "Lnz/etree do not match the pattern");
  std::vector<double> ax(nnz);
  for (int64_t q = 0; q < nnz; ++q) ax[q] = ValueIO<T>::get(Ax[q]);
  try {
    E.solver->factor(ax.data(), false);
  } catch (const Error& e) {
    if (e.code == kErrNumeric) return -1;  // QDLDL: exact zero pivot (16-bit: also a value out of range)
    throw;
  }
  E.solver->export_factor(Lp64, Li, Lx, D, Dinv, prec);
  return E.solver->stats().positive_pivots;
}



template <class TF, class TX>
int solve_impl(int32_t n, const int64_t* Lp, const int32_t* Li, const TF* Lx, const TF* Dinv, TX* x) {
  static_assert(std::is_same<TX, double>::value || std::is_same<TX, float>::value, "x precision");
  const int prec = std::is_same<TF, double>::value ? kF64 : kF32;
  QdldlEngine& E = engines().solve[prec];
  if (!E.solver) E.solver = std::make_unique<GmresLdl>(prec, kF64, kF64);
  E.solver->load_factor_csc(n, Lp, Li, Lx, Dinv);
  DevBuf xd(sizeof(double) * n);
  std::vector<double> xh(x, x + n);
  hipStream_t st = E.solver->stream();
  MPIR_HIP(hipMemcpyAsync(xd.get(), xh.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  E.solver->precond_apply_dev(xd.as<double>());
  MPIR_HIP(hipMemcpyAsync(xh.data(), xd.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
  E.solver->sync_check();
  for (int32_t i = 0; i < n; ++i) x[i] = static_cast<TX>(xh[i]);
  return 0;
}

// LUIR_QDLDL<FACTPREC, double, double, double, INT> (luir.hpp:33-241): the
// matrix as given (natural order, no scaling), factored on the GPU, then
// classic iterative refinement on the GPU.
template <class T>
int luir_impl(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* b, double* x,
              mpir_luir_stats* stats, int maxiter) {
  MPIR_CHECK(n >= 1 && Ap && Ai && Ax && b && x, "bad LUIR arguments");
  const int prec = std::is_same<T, double>::value ? kF64 : kF32;
  QdldlEngine& E = engines().luir[prec];
  if (!E.solver) {
    E.solver = std::make_unique<GmresLdl>(prec, kF64, kF64);
    E.solver->set_ordering(kOrderNatural);
    E.solver->set_scaling(kScaleNone);
  }
  const auto ts = std::chrono::steady_clock::now();
  // QDLDL_etree's checks and messages (luir.hpp:86-102 report and exit)
  std::vector<int32_t> work(n), lnz(n), par(n);
  const int64_t sum = etree_counts(n, Ap, Ai, work.data(), lnz.data(), par.data(), true);
  if (sum == -1) {
    std::fprintf(stderr, "!!! The input matrix is not in upper triangular form\n");
    return -1;
  }
  if (sum == -2) {
    std::fprintf(stderr, "!!! total nonzeros in L below diagonal overflows in INT\n");
    return -2;
  }
  const int64_t nnz = Ap[n];
  if (!E.same_pattern(n, Ap, Ai)) E.analyze(n, Ap, Ai);
  const double t_sym = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
  const auto tf = std::chrono::steady_clock::now();
  try {
    E.solver->factor(Ax, false);
  } catch (const Error& e) {
    if (e.code == kErrNumeric) return -1;  // QDLDL: exact zero pivot
    throw;
  }
  const double t_fact = std::chrono::duration<double>(std::chrono::steady_clock::now() - tf).count();
  LuirResult r;
  E.solver->luir(b, x, maxiter, &r);
  if (stats) {
    stats->n = n;
    stats->nnz = nnz;
    stats->lnz = sum;
    stats->num_refinement_steps = maxiter == 0 ? 0 : r.steps;
    stats->relative_residual = r.rel_residual;
    stats->time_sym = t_sym;
    stats->time_num_fact = t_fact;
    stats->time_solve = r.time_solve * 1e-3;
    stats->time_ref = r.time_ref * 1e-3;
  }
  return 0;
}

int spmv_upper_impl(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x,
                    const double* b, double* y) {
  MPIR_CHECK(n >= 1 && Ap && Ai && Ax && x && y, "bad SpMV arguments");
  CscUpper A;
  A.n = n;
  A.p.assign(Ap, Ap + n + 1);
  A.i.assign(Ai, Ai + Ap[n]);
  A.x.assign(Ax, Ax + Ap[n]);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) MPIR_CHECK(Ai[q] >= 0 && Ai[q] <= j, "input must be upper triangular");
  CsrFull F = csr_from_upper(A);
  hipStream_t st;
  MPIR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBuf dp(sizeof(int32_t) * (n + 1)), dj(sizeof(int32_t) * std::max<size_t>(1, F.j.size())),
      dx(sizeof(double) * std::max<size_t>(1, F.x.size())), dv(sizeof(double) * n), dy(sizeof(double) * n),
      db(sizeof(double) * n);
  MPIR_HIP(hipMemcpyAsync(dp.get(), F.p.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
  MPIR_HIP(hipMemcpyAsync(dj.get(), F.j.data(), sizeof(int32_t) * F.j.size(), hipMemcpyHostToDevice, st));
  MPIR_HIP(hipMemcpyAsync(dx.get(), F.x.data(), sizeof(double) * F.x.size(), hipMemcpyHostToDevice, st));
  MPIR_HIP(hipMemcpyAsync(dv.get(), x, sizeof(double) * n, hipMemcpyHostToDevice, st));
  CsrDev D{n, dp.as<int32_t>(), dj.as<int32_t>(), dx.as<double>()};
  if (b) {
    MPIR_HIP(hipMemcpyAsync(db.get(), b, sizeof(double) * n, hipMemcpyHostToDevice, st));
    residual_f64(D, dv.as<double>(), db.as<double>(), dy.as<double>(), st);
  } else {
    spmv_f64(D, dv.as<double>(), dy.as<double>(), st);
  }
  MPIR_HIP(hipMemcpyAsync(y, dy.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
  MPIR_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace
}  // namespace mpir

using namespace mpir;

struct mpir_gmres_ldl {
  std::unique_ptr<GmresLdl> s;
  std::vector<int32_t> Ap, Ai;
};
struct mpir_server {
  std::unique_ptr<mpir::Server> s;
};

namespace {
// symbolic phases of every pattern seen by the batch entry points
mpir::PatternCache& pattern_cache() {
  static mpir::PatternCache cache;
  return cache;
}
}  // namespace

// The int32 / int64 forms of the 16-bit QDLDL entry points (the int64 form
// narrows every index to int32, as the other int64 entry points do).
template <class B>
int32_t factor16_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const B* Ax, int32_t* Lp, int32_t* Li, B* Lx,
                     B* D, B* Dinv, const int32_t* Lnz, const int32_t* etree) {
  int64_t ret = 0;
  const int rc = guarded([&] {
    MPIR_CHECK(n >= 1 && Ap && Ai && Ax && Lp && Li && Lx && D && Dinv && Lnz && etree, "bad factor arguments");
    std::vector<int64_t> lp(static_cast<size_t>(n) + 1);
    ret = factor_impl<B>(n, Ap, Ai, Ax, lp.data(), Li, Lx, D, Dinv, Lnz, etree);
    for (int32_t k = 0; k <= n; ++k) Lp[k] = static_cast<int32_t>(lp[k]);
    return 0;
  });
  return rc != 0 ? rc : static_cast<int32_t>(ret);
}
template <class B>
int64_t factor16_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const B* Ax, int64_t* Lp, int64_t* Li, B* Lx,
                     B* D, B* Dinv, const int64_t* Lnz, const int64_t* etree) {
  int64_t ret = 0;
  const int rc = guarded([&] {
    int32_t n32;
    MPIR_CHECK(narrow(n, &n32) && n32 >= 1 && Ap && Ai && Ax && Lp && Li && Lx && D && Dinv && Lnz && etree,
               "bad factor arguments");
    std::vector<int32_t> p = narrow_vec(Ap, n + 1), i = narrow_vec(Ai, Ap[n]);
    std::vector<int32_t> l = narrow_vec(Lnz, n), e = narrow_vec(etree, n);
    int64_t lnz = 0;
    for (int64_t k = 0; k < n; ++k) lnz += Lnz[k];
    std::vector<int32_t> li(static_cast<size_t>(std::max<int64_t>(lnz, 1)));
    ret = factor_impl<B>(n32, p.data(), i.data(), Ax, Lp, li.data(), Lx, D, Dinv, l.data(), e.data());
    for (int64_t k = 0; k < lnz; ++k) Li[k] = li[k];
    return 0;
  });
  return rc != 0 ? rc : ret;
}
// QDLDL_solve with a 16-bit factor: its values are exact in FP32, so the
// solve runs on the FP32 engine with the arithmetic of XPREC.
template <int P, class TX>
int solve16(int32_t n, const int64_t* Lp, const int32_t* Li, const uint16_t* Lx, const uint16_t* Dinv, TX* x) {
  MPIR_CHECK(n >= 1 && Lp && Li && Lx && Dinv && x, "bad solve arguments");
  const int64_t nnz = Lp[n];
  std::vector<float> lx(static_cast<size_t>(std::max<int64_t>(nnz, 1))), di(n);
  for (int64_t q = 0; q < nnz; ++q) lx[q] = static_cast<float>(ValueIO<Bits16<P>>::get(Bits16<P>{Lx[q]}));
  for (int32_t i = 0; i < n; ++i) di[i] = static_cast<float>(ValueIO<Bits16<P>>::get(Bits16<P>{Dinv[i]}));
  return solve_impl<float, TX>(n, Lp, Li, lx.data(), di.data(), x);
}
template <int P, class TX>
int32_t solve16_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx, const uint16_t* Dinv, TX* x) {
  return guarded([&] {
    MPIR_CHECK(n >= 1 && Lp, "bad solve arguments");
    std::vector<int64_t> lp(Lp, Lp + n + 1);
    return solve16<P, TX>(n, lp.data(), Li, Lx, Dinv, x);
  });
}
template <int P, class TX>
int64_t solve16_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx, const uint16_t* Dinv, TX* x) {
  return guarded([&] {
    int32_t n32;
    MPIR_CHECK(narrow(n, &n32) && n32 >= 1 && Lp && Li, "bad solve arguments");
    std::vector<int32_t> li = narrow_vec(Li, Lp[n]);
    return solve16<P, TX>(n32, Lp, li.data(), Lx, Dinv, x);
  });
}

extern "C" {

const char* mpir_last_error(void) { return g_last_error.c_str(); }
const char* mpir_version(void) { return "mpir_hip 0.1 (gfx950)"; }

int mpir_read_mtx_upper(const char* path, int32_t* n, int64_t* nnz, int32_t* symmetric, int32_t* Ap, int32_t* Ai,
                        double* Ax) {
  return guarded([&] {
    MPIR_CHECK(path && n && nnz, "bad Matrix Market arguments");
    MtxUpper M;
    try {
      M = read_mtx_upper(path);
    } catch (const std::invalid_argument& e) {
      throw Error(kErrInvalid, e.what());
    }
    *n = M.n;
    *nnz = static_cast<int64_t>(M.i.size());
    if (symmetric) *symmetric = M.symmetric ? 1 : 0;
    if (Ap) std::copy(M.p.begin(), M.p.end(), Ap);
    if (Ai) std::copy(M.i.begin(), M.i.end(), Ai);
    if (Ax) std::copy(M.x.begin(), M.x.end(), Ax);
    return 0;
  });
}

int64_t mpir_gmres_ldl_log(mpir_gmres_ldl_t h, char* buf, int64_t cap) {
  int64_t len = 0;
  const int rc = guarded([&] {
    MPIR_CHECK(h && h->s, "null solver handle");
    const std::string j = h->s->log_json();
    len = static_cast<int64_t>(j.size());
    if (buf && cap > 0) {
      const int64_t k = std::min<int64_t>(cap - 1, len);
      std::memcpy(buf, j.data(), static_cast<size_t>(k));
      buf[k] = 0;
    }
    return 0;
  });
  return rc < 0 ? rc : len;
}

int mpir_device_ok(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int32_t mpir_qdldl_etree_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* work, int32_t* Lnz,
                             int32_t* etree) {
  int64_t total = -1;
  const int rc = guarded([&] { return etree_impl(n, Ap, Ai, work, Lnz, etree, true, &total); });
  return rc != 0 ? rc : static_cast<int32_t>(total);
}

int64_t mpir_qdldl_etree_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, int64_t* work, int64_t* Lnz,
                             int64_t* etree) {
  int64_t total = -1;
  const int rc = guarded([&] {
    int32_t n32;
    MPIR_CHECK(narrow(n, &n32) && n32 >= 0, "n does not fit in int32");
    std::vector<int32_t> p = narrow_vec(Ap, n + 1);
    std::vector<int32_t> i = narrow_vec(Ai, Ap[n]);
    std::vector<int32_t> w(n32), l(n32), e(n32);
    etree_impl(n32, p.data(), i.data(), w.data(), l.data(), e.data(), false, &total);
    for (int64_t k = 0; k < n; ++k) {
      work[k] = w[k];
      Lnz[k] = l[k];
      etree[k] = e[k];
    }
    return 0;
  });
  return rc != 0 ? rc : total;
}

#define MPIR_FACTOR_I32(SUF, T)                                                                                       \
  int32_t mpir_qdldl_factor_##SUF##_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const T* Ax, int32_t* Lp,   \
                                        int32_t* Li, T* Lx, T* D, T* Dinv, const int32_t* Lnz, const int32_t* etree, \
                                        unsigned char*, int32_t*, T*) {                                                \
    int64_t ret = 0;                                                                                                   \
    const int rc = guarded([&] {                                                                                       \
      std::vector<int64_t> lp(static_cast<size_t>(n) + 1);                                                           \
      ret = factor_impl<T>(n, Ap, Ai, Ax, lp.data(), Li, Lx, D, Dinv, Lnz, etree);                                     \
      for (int32_t k = 0; k <= n; ++k) Lp[k] = static_cast<int32_t>(lp[k]);                                          \
      return 0;                                                                                                        \
    });                                                                                                                \
    return rc != 0 ? rc : static_cast<int32_t>(ret);                                                                  \
  }

#define MPIR_FACTOR_I64(SUF, T)                                                                                       \
  int64_t mpir_qdldl_factor_##SUF##_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const T* Ax, int64_t* Lp,   \
                                        int64_t* Li, T* Lx, T* D, T* Dinv, const int64_t* Lnz, const int64_t* etree, \
                                        unsigned char*, int64_t*, T*) {                                                \
    int64_t ret = 0;                                                                                                   \
    const int rc = guarded([&] {                                                                                       \
      int32_t n32;                                                                                                     \
      MPIR_CHECK(narrow(n, &n32) && n32 >= 1, "n does not fit in int32");                                             \
      std::vector<int32_t> p = narrow_vec(Ap, n + 1), i = narrow_vec(Ai, Ap[n]);                                     \
      std::vector<int32_t> l = narrow_vec(Lnz, n), e = narrow_vec(etree, n);                                         \
      int64_t lnz = 0;                                                                                                 \
      for (int64_t k = 0; k < n; ++k) lnz += Lnz[k];                                                                 \
      std::vector<int32_t> li(static_cast<size_t>(std::max<int64_t>(lnz, 1)));                                       \
      ret = factor_impl<T>(n32, p.data(), i.data(), Ax, Lp, li.data(), Lx, D, Dinv, l.data(), e.data());              \
      for (int64_t k = 0; k < lnz; ++k) Li[k] = li[k];                                                               \
      return 0;                                                                                                        \
    });                                                                                                                \
    return rc != 0 ? rc : ret;                                                                                         \
  }

MPIR_FACTOR_I32(f64, double)
MPIR_FACTOR_I32(f32, float)
MPIR_FACTOR_I64(f64, double)
MPIR_FACTOR_I64(f32, float)

// 16-bit factors and solves: the values are IEEE binary16 / bfloat16 bit
// patterns
#define MPIR_FACTOR16(SUF, P)                                                                                         \
  int32_t mpir_qdldl_factor_##SUF##_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const uint16_t* Ax,          \
                                        int32_t* Lp, int32_t* Li, uint16_t* Lx, uint16_t* D, uint16_t* Dinv,          \
                                        const int32_t* Lnz, const int32_t* etree, unsigned char*, int32_t*,          \
                                        uint16_t*) {                                                                   \
    using B = Bits16<P>;                                                                                              \
    return factor16_i32<B>(n, Ap, Ai, reinterpret_cast<const B*>(Ax), Lp, Li, reinterpret_cast<B*>(Lx),             \
                           reinterpret_cast<B*>(D), reinterpret_cast<B*>(Dinv), Lnz, etree);                         \
  }                                                                                                                    \
  int64_t mpir_qdldl_factor_##SUF##_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const uint16_t* Ax,          \
                                        int64_t* Lp, int64_t* Li, uint16_t* Lx, uint16_t* D, uint16_t* Dinv,          \
                                        const int64_t* Lnz, const int64_t* etree, unsigned char*, int64_t*,          \
                                        uint16_t*) {                                                                   \
    using B = Bits16<P>;                                                                                              \
    return factor16_i64<B>(n, Ap, Ai, reinterpret_cast<const B*>(Ax), Lp, Li, reinterpret_cast<B*>(Lx),             \
                           reinterpret_cast<B*>(D), reinterpret_cast<B*>(Dinv), Lnz, etree);                         \
  }                                                                                                                    \
  int32_t mpir_qdldl_solve_##SUF##_f32_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx,       \
                                           const uint16_t* Dinv, float* x) {                                          \
    return solve16_i32<P, float>(n, Lp, Li, Lx, Dinv, x);                                                            \
  }                                                                                                                    \
  int32_t mpir_qdldl_solve_##SUF##_f64_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const uint16_t* Lx,       \
                                           const uint16_t* Dinv, double* x) {                                         \
    return solve16_i32<P, double>(n, Lp, Li, Lx, Dinv, x);                                                           \
  }                                                                                                                    \
  int64_t mpir_qdldl_solve_##SUF##_f32_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx,       \
                                           const uint16_t* Dinv, float* x) {                                          \
    return solve16_i64<P, float>(n, Lp, Li, Lx, Dinv, x);                                                            \
  }                                                                                                                    \
  int64_t mpir_qdldl_solve_##SUF##_f64_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const uint16_t* Lx,       \
                                           const uint16_t* Dinv, double* x) {                                         \
    return solve16_i64<P, double>(n, Lp, Li, Lx, Dinv, x);                                                           \
  }
MPIR_FACTOR16(f16, kF16)
MPIR_FACTOR16(bf16, kBF16)

#define MPIR_SOLVE_I32(SUF, TF, TX)                                                                                  \
  int32_t mpir_qdldl_solve_##SUF##_i32(int32_t n, const int32_t* Lp, const int32_t* Li, const TF* Lx,               \
                                       const TF* Dinv, TX* x) {                                                      \
    return guarded([&] {                                                                                             \
      MPIR_CHECK(n >= 1 && Lp && Li && Lx && Dinv && x, "bad solve arguments");                                     \
      std::vector<int64_t> lp(Lp, Lp + n + 1);                                                                       \
      return solve_impl<TF, TX>(n, lp.data(), Li, Lx, Dinv, x);                                                      \
    });                                                                                                              \
  }
#define MPIR_SOLVE_I64(SUF, TF, TX)                                                                                  \
  int64_t mpir_qdldl_solve_##SUF##_i64(int64_t n, const int64_t* Lp, const int64_t* Li, const TF* Lx,               \
                                       const TF* Dinv, TX* x) {                                                      \
    return guarded([&] {                                                                                             \
      int32_t n32;                                                                                                   \
      MPIR_CHECK(narrow(n, &n32) && n32 >= 1 && Lp && Li && Lx && Dinv && x, "bad solve arguments");                \
      std::vector<int32_t> li = narrow_vec(Li, Lp[n]);                                                              \
      return solve_impl<TF, TX>(n32, Lp, li.data(), Lx, Dinv, x);                                                    \
    });                                                                                                              \
  }
#define MPIR_LUIR(SUF, TF)                                                                                     \
  int32_t mpir_luir_qdldl_##SUF##_i32(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax,          \
                                      const double* b, double* x, mpir_luir_stats* stats, int32_t maxiter) {      \
    return guarded([&] { return luir_impl<TF>(n, Ap, Ai, Ax, b, x, stats, maxiter); });                           \
  }                                                                                                              \
  int64_t mpir_luir_qdldl_##SUF##_i64(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax,          \
                                      const double* b, double* x, mpir_luir_stats* stats, int32_t maxiter) {      \
    return guarded([&] {                                                                                         \
      int32_t n32;                                                                                               \
      MPIR_CHECK(narrow(n, &n32) && n32 >= 1 && Ap && Ai, "bad LUIR arguments");                                 \
      std::vector<int32_t> p = narrow_vec(Ap, n + 1);                                                            \
      std::vector<int32_t> i = narrow_vec(Ai, Ap[n]);                                                            \
      return luir_impl<TF>(n32, p.data(), i.data(), Ax, b, x, stats, maxiter);                                    \
    });                                                                                                          \
  }
MPIR_LUIR(f64, double)
MPIR_LUIR(f32, float)

MPIR_SOLVE_I32(f64_f64, double, double)
MPIR_SOLVE_I32(f32_f64, float, double)
MPIR_SOLVE_I32(f32_f32, float, float)
MPIR_SOLVE_I64(f64_f64, double, double)
MPIR_SOLVE_I64(f32_f64, float, double)
MPIR_SOLVE_I64(f32_f32, float, float)

int mpir_gmres_ldl_create(int uf, int uw, int ur, mpir_gmres_ldl_t* out) {
  return guarded([&] {
    MPIR_CHECK(out, "null output handle");
    *out = nullptr;
    auto h = std::make_unique<mpir_gmres_ldl>();
    h->s = std::make_unique<GmresLdl>(uf, uw, ur);
    *out = h.release();
    return 0;
  });
}

void mpir_gmres_ldl_destroy(mpir_gmres_ldl_t h) { delete h; }

#define MPIR_HANDLE_CHECK \
  MPIR_CHECK(h && h->s, "null solver handle")

int mpir_gmres_ldl_set_max_ir(mpir_gmres_ldl_t h, int64_t n) {
  return guarded([&] { MPIR_HANDLE_CHECK; MPIR_CHECK(n >= 0, "negative iteration count"); h->s->set_max_ir(n); return 0; });
}
int mpir_gmres_ldl_set_max_gmres(mpir_gmres_ldl_t h, int64_t n) {
  return guarded([&] { MPIR_HANDLE_CHECK; MPIR_CHECK(n >= 0, "negative iteration count"); h->s->set_max_gmres(n); return 0; });
}
int mpir_gmres_ldl_set_tol(mpir_gmres_ldl_t h, double tol) {
  return guarded([&] { MPIR_HANDLE_CHECK; h->s->set_tol(tol); return 0; });
}
int mpir_gmres_ldl_set_scaling(mpir_gmres_ldl_t h, int mode) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_scaling(mode);
    return 0;
  });
}
int mpir_gmres_ldl_set_preconditioner(mpir_gmres_ldl_t h, int kind, int32_t ilu_k, int32_t max_sweeps) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_preconditioner(kind, ilu_k, max_sweeps);
    return 0;
  });
}
int mpir_gmres_ldl_ilu_nnz(mpir_gmres_ldl_t h, int64_t* nnz) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(nnz, "null output");
    MPIR_CHECK(h->s->analyzed() && h->s->preconditioner() == kPrecIlu, "no FGPILU pattern analyzed");
    *nnz = h->s->ilu_nnz();
    return 0;
  });
}
int mpir_gmres_ldl_export_ilu(mpir_gmres_ldl_t h, int32_t* Up, int32_t* Ui, double* Ux, double* Lx, double* UDinv) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->export_ilu(Up, Ui, Ux, Lx, UDinv);
    return 0;
  });
}
int mpir_gmres_ldl_set_schedule(mpir_gmres_ldl_t h, int schedule) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_schedule(schedule);
    return 0;
  });
}
int mpir_gmres_ldl_set_amalgamation(mpir_gmres_ldl_t h, int level) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_relax(level);
    return 0;
  });
}
int mpir_gmres_ldl_set_panel_split(mpir_gmres_ldl_t h, int min_tasks) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_panel_split(min_tasks);
    return 0;
  });
}
int mpir_gmres_ldl_set_factor_schedule(mpir_gmres_ldl_t h, int schedule) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_factor_schedule(schedule);
    return 0;
  });
}
int mpir_gmres_ldl_factor_schedule_info(mpir_gmres_ldl_t h, int* schedule, int32_t* dag_level, int32_t* nlevels,
                                        int64_t* dag_tasks) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    int32_t lv = 0, nl = 0;
    int64_t nt = 0;
    h->s->factor_schedule_info(&lv, &nl, &nt);
    if (schedule) *schedule = h->s->factor_schedule();
    if (dag_level) *dag_level = lv;
    if (nlevels) *nlevels = nl;
    if (dag_tasks) *dag_tasks = nt;
    return 0;
  });
}

int mpir_gmres_ldl_set_ordering(mpir_gmres_ldl_t h, int ordering) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(ordering == MPIR_ORDER_NATURAL || ordering == MPIR_ORDER_AMD || ordering == MPIR_ORDER_ND,
               "unknown ordering");
    h->s->set_ordering(ordering);
    return 0;
  });
}
int mpir_gmres_ldl_analyze(mpir_gmres_ldl_t h, int32_t n, const int32_t* Ap, const int32_t* Ai) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->analyze(n, Ap, Ai);
    h->Ap.assign(Ap, Ap + n + 1);
    h->Ai.assign(Ai, Ai + Ap[n]);
    return 0;
  });
}
int mpir_gmres_ldl_factor(mpir_gmres_ldl_t h, const double* Ax) {
  return guarded([&] { MPIR_HANDLE_CHECK; MPIR_CHECK(Ax, "null values"); h->s->factor(Ax, false); return 0; });
}
int mpir_gmres_ldl_factor_dev(mpir_gmres_ldl_t h, const double* Ax) {
  return guarded([&] { MPIR_HANDLE_CHECK; MPIR_CHECK(Ax, "null values"); h->s->factor(Ax, true); return 0; });
}
int mpir_gmres_ldl_compute(mpir_gmres_ldl_t h, int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(n >= 1 && Ap && Ai && Ax, "bad compute arguments");
    const bool same = h->s->analyzed() && static_cast<int32_t>(h->Ap.size()) == n + 1 &&
                      std::equal(h->Ap.begin(), h->Ap.end(), Ap) && static_cast<int64_t>(h->Ai.size()) == Ap[n] &&
                      std::equal(h->Ai.begin(), h->Ai.end(), Ai);
    if (!same) {
      h->s->analyze(n, Ap, Ai);
      h->Ap.assign(Ap, Ap + n + 1);
      h->Ai.assign(Ai, Ai + Ap[n]);
    }
    h->s->factor(Ax, false);
    return 0;
  });
}
int mpir_gmres_ldl_solve(mpir_gmres_ldl_t h, const double* b, double* x, int32_t nrhs, int64_t ld) {
  return guarded([&] { MPIR_HANDLE_CHECK; MPIR_CHECK(b && x, "null vectors"); h->s->solve(b, x, nrhs, ld, false); return 0; });
}
int mpir_gmres_ldl_solve_dev(mpir_gmres_ldl_t h, const double* b, double* x, int32_t nrhs, int64_t ld) {
  return guarded([&] { MPIR_HANDLE_CHECK; MPIR_CHECK(b && x, "null vectors"); h->s->solve(b, x, nrhs, ld, true); return 0; });
}
// ---------------------------------------------------------------- batches
int mpir_batch_analyze(mpir_gmres_ldl_t h, int32_t nsys, const int32_t* n, const int32_t* const* Ap,
                       const int32_t* const* Ai, int32_t threads) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(nsys >= 1 && n && Ap && Ai, "bad batch description");
    std::vector<int32_t> nv(n, n + nsys);
    std::vector<const int32_t*> ap(Ap, Ap + nsys), ai(Ai, Ai + nsys);
    std::vector<std::shared_ptr<const PatternEntry>> parts;
    try {
      parts = pattern_cache().get_many(nv, ap, ai, h->s->ordering(), h->s->relax(), std::max(1, threads));
    } catch (const std::invalid_argument& e) {
      throw Error(kErrInvalid, e.what());
    }
    h->s->analyze_batch(parts);
    return 0;
  });
}
int mpir_batch_factor(mpir_gmres_ldl_t h, const double* const* Ax) {
  return guarded([&] { MPIR_HANDLE_CHECK; h->s->factor_batch(Ax, false); return 0; });
}
int mpir_batch_factor_dev(mpir_gmres_ldl_t h, const double* const* Ax) {
  return guarded([&] { MPIR_HANDLE_CHECK; h->s->factor_batch(Ax, true); return 0; });
}
int mpir_batch_solve(mpir_gmres_ldl_t h, const int32_t* nrhs, const double* const* b, double* const* x) {
  return guarded([&] { MPIR_HANDLE_CHECK; h->s->solve_batch(nrhs, b, x, false); return 0; });
}
int mpir_batch_solve_dev(mpir_gmres_ldl_t h, const int32_t* nrhs, const double* const* b, double* const* x) {
  return guarded([&] { MPIR_HANDLE_CHECK; h->s->solve_batch(nrhs, b, x, true); return 0; });
}
int mpir_batch_stats(mpir_gmres_ldl_t h, int32_t* max_rhs, int32_t* positive_pivots, int32_t* ir_iterations,
                     int32_t* gmres_iterations, int32_t* stop_reason, double* final_rel_residual, double* done_ms) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(h->s->is_batch(), "handle holds no batch");
    const BatchStats& b = h->s->batch_stats();
    if (max_rhs) *max_rhs = b.max_rhs;
    auto put = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(positive_pivots, b.positive_pivots);
    put(ir_iterations, b.ir_iterations);
    put(gmres_iterations, b.gmres_iterations);
    put(stop_reason, b.stop_reason);
    put(final_rel_residual, b.final_rel_residual);
    put(done_ms, b.done_ms);
    return 0;
  });
}
int mpir_batch_graph(mpir_gmres_ldl_t h, int enable, int32_t* counts) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    if (enable >= 0) h->s->set_batch_graph(enable != 0);
    if (counts) {
      const BatchStats& b = h->s->batch_stats();
      counts[0] = b.graph_nodes;
      counts[1] = b.arnoldi_steps;
      counts[2] = b.precond_sweeps;
    }
    return 0;
  });
}
int mpir_qdldl_solve_plan_builds(int64_t* builds) {
  return guarded([&] {
    MPIR_CHECK(builds, "null output");
    int64_t b = 0;
    for (const QdldlEngine& E : engines().solve)
      if (E.solver) b += E.solver->plan_builds();
    *builds = b;
    return 0;
  });
}
int mpir_pattern_cache_info(int64_t* entries, int64_t* hits, int64_t* misses) {
  return guarded([&] {
    if (entries) *entries = static_cast<int64_t>(pattern_cache().size());
    if (hits) *hits = pattern_cache().hits();
    if (misses) *misses = pattern_cache().misses();
    return 0;
  });
}
int mpir_batch_plan_info(int32_t nsys, const int32_t* n, const int32_t* const* Ap, const int32_t* const* Ai,
                         int ordering, int64_t* info) {
  return guarded([&] {
    MPIR_CHECK(nsys >= 1 && n && Ap && Ai && info, "bad batch description");
    MPIR_CHECK(ordering >= 0 && ordering <= 2, "ordering must be 0, 1 or 2");
    std::vector<PatternEntry> own;
    own.reserve(nsys);
    std::vector<const PatternEntry*> raw;
    try {
      for (int32_t s = 0; s < nsys; ++s) own.push_back(analyze_entry(n[s], Ap[s], Ai[s], ordering, kDefaultAmalgamation));
    } catch (const std::invalid_argument& e) {
      throw Error(kErrInvalid, e.what());
    }
    for (const auto& e : own) raw.push_back(&e);
    BatchSystem B;
    FactorPlan P;
    combine_patterns(raw, B, P);
    info[0] = B.col_base.back();
    info[1] = B.nnz_l;
    info[2] = static_cast<int64_t>(B.flops);
    info[3] = P.sn.ns;
    info[4] = P.sn.nlevels;
    info[5] = P.panel_entries;
    info[6] = static_cast<int64_t>(P.pan_task.size());
    info[7] = static_cast<int64_t>(P.dag_fwd.size());
    // per system: nnz(L) and the footprint estimate
    for (int32_t s = 0; s < nsys; ++s) {
      info[8 + 2 * s] = own[s].an.nnz_l;
      info[9 + 2 * s] = own[s].est_bytes;
    }
    return 0;
  });
}

// ---------------------------------------------------------------- serving
int mpir_server_create(int uf, int lanes, int64_t budget_bytes, int32_t max_batch, double tol, int ordering,
                       int32_t keep_x, int32_t threads, mpir_server_t* out) {
  return guarded([&] {
    MPIR_CHECK(out, "null output handle");
    *out = nullptr;
    ServeOptions o;
    o.uf = uf;
    o.lanes = lanes;
    o.budget_bytes = budget_bytes;
    o.max_batch = max_batch;
    o.tol = tol;
    o.ordering = ordering;
    o.keep_x = keep_x;
    o.threads = std::max(1, threads);
    auto h = std::make_unique<mpir_server>();
    h->s = std::make_unique<Server>(o);
    *out = h.release();
    return 0;
  });
}
void mpir_server_destroy(mpir_server_t h) { delete h; }
int mpir_server_add_pattern(mpir_server_t h, int32_t n, const int32_t* Ap, const int32_t* Ai, const double* base,
                            int32_t n1) {
  return guarded([&] {
    MPIR_CHECK(h && h->s, "null server handle");
    return h->s->add_pattern(n, Ap, Ai, base, n1);
  });
}
int mpir_server_add_patterns(mpir_server_t h, int32_t count, const int32_t* n, const int32_t* const* Ap,
                             const int32_t* const* Ai, const double* const* base, const int32_t* n1, int32_t threads) {
  return guarded([&] {
    MPIR_CHECK(h && h->s, "null server handle");
    return h->s->add_patterns(count, n, Ap, Ai, base, n1, std::max(1, threads));
  });
}
int mpir_server_run(mpir_server_t h, int64_t nreq, const int32_t* pattern, const uint64_t* seed, const int32_t* nrhs,
                    const double* t_arrive, double* t_admit, double* t_done, int32_t* batch, int32_t* lane,
                    int32_t* status, int32_t* ir_max, int32_t* gmres_max, double* relres_max) {
  return guarded([&] {
    MPIR_CHECK(h && h->s, "null server handle");
    MPIR_CHECK(nreq >= 0 && (nreq == 0 || (pattern && seed && nrhs && t_arrive)), "bad request arrays");
    std::vector<ServeRecord> rs(nreq);
    for (int64_t i = 0; i < nreq; ++i) {
      rs[i].pattern = pattern[i];
      rs[i].seed = seed[i];
      rs[i].nrhs = nrhs[i];
      rs[i].t_arrive = t_arrive[i];
    }
    h->s->run(rs);
    for (int64_t i = 0; i < nreq; ++i) {
      if (t_admit) t_admit[i] = rs[i].t_admit;
      if (t_done) t_done[i] = rs[i].t_done;
      if (batch) batch[i] = rs[i].batch;
      if (lane) lane[i] = rs[i].lane;
      if (status) status[i] = rs[i].status;
      if (ir_max) ir_max[i] = rs[i].ir_max;
      if (gmres_max) gmres_max[i] = rs[i].gmres_max;
      if (relres_max) relres_max[i] = rs[i].rel_residual_max;
    }
    return 0;
  });
}
int mpir_server_submit(mpir_server_t h, int32_t pattern, const double* Ax, const double* b, int32_t nrhs,
                       int64_t* id) {
  return guarded([&] {
    MPIR_CHECK(h && h->s && id, "null argument");
    *id = h->s->submit(pattern, Ax, b, nrhs);
    return 0;
  });
}
int mpir_server_poll(mpir_server_t h, int64_t id, double timeout_ms, double* x, mpir_server_request* info,
                     int32_t release) {
  return guarded([&] {
    MPIR_CHECK(h && h->s, "null server handle");
    ServeRequest r;
    if (!h->s->wait(id, timeout_ms, &r, x, release != 0)) return 0;
    if (info) {
      info->status = r.status;
      info->batch = r.batch;
      info->lane = r.lane;
      info->ir_max = r.ir_max;
      info->gmres_max = r.gmres_max;
      info->rel_residual_max = r.rel_residual_max;
      info->queue_ms = r.t_admit - r.t_arrive;
      info->latency_ms = r.t_done - r.t_arrive;
    }
    if (r.status < 0) {
      set_last_error("request " + std::to_string(id) + " failed in batch " + std::to_string(r.batch));
      return r.status;
    }
    return 1;
  });
}
int mpir_server_plan_phases(mpir_server_t h, double* join_ms, double* upload_ms, int64_t* regrowths) {
  return guarded([&] {
    MPIR_CHECK(h && h->s, "null server handle");
    const ServeSummary m = h->s->summary();
    if (join_ms) *join_ms = m.plan_join_ms;
    if (upload_ms) *upload_ms = m.plan_upload_ms;
    if (regrowths) *regrowths = m.regrowths;
    return 0;
  });
}
int mpir_server_summary(mpir_server_t h, int64_t* counts, double* wall_ms) {
  return guarded([&] {
    MPIR_CHECK(h && h->s, "null server handle");
    const ServeSummary m = h->s->summary();
    if (wall_ms) {
      wall_ms[1] = m.prepare_ms;
      wall_ms[2] = m.plan_ms;
      wall_ms[3] = m.factor_ms;
      wall_ms[4] = m.solve_ms;
    }
    if (counts) {
      counts[0] = m.completed;
      counts[1] = m.failed;
      counts[2] = m.batches;
      counts[3] = m.max_backlog;
      counts[4] = m.peak_inflight_bytes;
      counts[5] = m.max_concurrent;
    }
    if (wall_ms) *wall_ms = m.wall_ms;
    return 0;
  });
}
int mpir_server_solution(mpir_server_t h, int64_t i, double* x) {
  return guarded([&] {
    MPIR_CHECK(h && h->s && x, "null argument");
    const std::vector<double>& v = h->s->solution(i);
    std::memcpy(x, v.data(), v.size() * sizeof(double));
    return 0;
  });
}
int mpir_server_request_data(mpir_server_t h, int32_t pattern, uint64_t seed, int32_t nrhs, double* Ax, double* b,
                             double* x_true) {
  return guarded([&] {
    MPIR_CHECK(h && h->s && Ax && b && x_true, "null argument");
    MPIR_CHECK(nrhs >= 1, "nrhs must be positive");
    materialize(h->s->pattern(pattern), seed, nrhs, Ax, b, x_true);
    return 0;
  });
}

int mpir_gmres_ldl_stats(mpir_gmres_ldl_t h, mpir_gmres_stats* out) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(out, "null stats");
    const SolveStats& s = h->s->stats();
    out->n = s.n;
    out->nnz = s.nnz;
    out->nnz_l = s.nnz_l;
    out->nsupernodes = s.nsupernodes;
    out->nlevels = s.nlevels;
    out->max_front = s.max_front;
    out->factor_flops = s.factor_flops;
    out->positive_pivots = s.positive_pivots;
    out->ir_iterations = s.ir_iterations;
    out->gmres_iterations = s.gmres_iterations;
    out->precond_applications = s.precond_applications;
    out->stop_reason = s.stop_reason;
    out->final_rel_residual = s.final_rel_residual;
    out->analyze_ms = s.analyze_ms;
    out->factor_ms = s.factor_ms;
    out->solve_ms = s.solve_ms;
    out->device_bytes = s.device_bytes;
    out->ruiz_sweeps = s.ruiz_sweeps;
    out->preconditioner = s.preconditioner;
    out->ilu_sweeps = s.ilu_sweeps;
    out->ilu_nl_residual0 = s.ilu_nl_residual0;
    out->ilu_nl_residual = s.ilu_nl_residual;
    return 0;
  });
}
int mpir_gmres_ldl_perm(mpir_gmres_ldl_t h, int32_t* perm) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(h->s->analyzed(), "no pattern analyzed");
    std::memcpy(perm, h->s->perm().data(), sizeof(int32_t) * h->s->n());
    return 0;
  });
}
int mpir_gmres_ldl_export_factor(mpir_gmres_ldl_t h, int64_t* Lp, int32_t* Li, double* Lx, double* D, double* Dinv) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(Lp, "null Lp");
    h->s->export_factor(Lp, Li, Lx, D, Dinv, kF64);
    return 0;
  });
}
int mpir_gmres_ldl_precond_dev(mpir_gmres_ldl_t h, double* x) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(h->s->factored(), "no factor");
    h->s->precond_apply_dev(x);
    h->s->sync_check();
    return 0;
  });
}
int mpir_gmres_ldl_precond_dev_noalloc(mpir_gmres_ldl_t h, double* x) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(h->s->factored(), "no factor");
    precond_apply_noalloc(*h->s, x);
    return 0;
  });
}
int mpir_gmres_ldl_precond_multi_dev(mpir_gmres_ldl_t h, double* x, int32_t nrhs) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(h->s->factored(), "no factor");
    MPIR_CHECK(x && nrhs >= 1, "bad right-hand sides");
    h->s->precond_apply_slots(x, nrhs);
    h->s->sync_check();
    return 0;
  });
}
int mpir_gmres_ldl_spmv_dev(mpir_gmres_ldl_t h, const double* x, double* y) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(h->s->factored(), "no factor");
    h->s->spmv_dev(x, y);
    return 0;
  });
}
int mpir_gmres_ldl_time(mpir_gmres_ldl_t h, int what, int reps, double* avg_ms) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(avg_ms, "null output");
    *avg_ms = h->s->time_kernels(what, reps);
    return 0;
  });
}
int mpir_gmres_ldl_trace_precond(mpir_gmres_ldl_t h, uint64_t* out, int64_t cap, int64_t* ntask, int64_t* nfwd) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(ntask && nfwd && (out || cap == 0) && cap >= 0, "bad trace buffer");
    *ntask = h->s->trace_precond(out, cap, nfwd);
    return 0;
  });
}
int mpir_gmres_ldl_supernode_info(mpir_gmres_ldl_t h, int32_t* level, int32_t* ncols, int32_t* nrows, int32_t cap,
                                  int32_t* ns) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(ns && cap >= 0 && (cap == 0 || (level && ncols && nrows)), "bad supernode buffers");
    *ns = h->s->supernode_info(level, ncols, nrows, cap);
    return 0;
  });
}
int mpir_gmres_ldl_trace_factor(mpir_gmres_ldl_t h, uint64_t* out, int64_t cap, int64_t* ntask) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    MPIR_CHECK(ntask && (out || cap == 0) && cap >= 0, "bad trace buffer");
    *ntask = h->s->trace_factor(out, cap);
    return 0;
  });
}
int mpir_gmres_ldl_set_tuning(mpir_gmres_ldl_t h, int32_t upd_group, int32_t super_tiles, int32_t narrow_rows) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    h->s->set_tuning(FactorTuning{upd_group, super_tiles}, SolveTuning{narrow_rows});
    return 0;
  });
}
int mpir_gmres_ldl_autotune(mpir_gmres_ldl_t h, int reps, int32_t* choice, double* ms) {
  return guarded([&] {
    MPIR_HANDLE_CHECK;
    const AutotuneResult r = h->s->autotune(reps);
    if (choice) {
      choice[0] = r.factor.upd_group;
      choice[1] = r.factor.super_tiles;
      choice[2] = r.solve.narrow_rows;
      choice[3] = r.schedule;
      choice[4] = r.candidates;
    }
    if (ms) {
      ms[0] = r.factor_ms;
      ms[1] = r.precond_ms;
      ms[2] = r.factor_ms_before;
      ms[3] = r.precond_ms_before;
    }
    return 0;
  });
}
void* mpir_gmres_ldl_stream(mpir_gmres_ldl_t h) { return h && h->s ? static_cast<void*>(h->s->stream()) : nullptr; }

int mpir_spmv_upper_f64(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x, double* y) {
  return guarded([&] { return spmv_upper_impl(n, Ap, Ai, Ax, x, nullptr, y); });
}
int mpir_residual_upper_f64(int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax, const double* x,
                            const double* b, double* r) {
  return guarded([&] {
    MPIR_CHECK(b, "null b");
    return spmv_upper_impl(n, Ap, Ai, Ax, x, b, r);
  });
}

int mpir_gen_kkt(int dim, int k, double h_shift, int64_t m, int a_mode, int cnt_lo, int cnt_hi, double d_value,
                 uint64_t seed, int32_t* n_out, int64_t* nnz_out, int32_t* n1_out, int32_t* Ap, int32_t* Ai,
                 double* Ax, double* x_true, double* b) {
  return guarded([&] {
    KktSpec s;
    s.dim = dim;
    s.k = k;
    s.h_shift = h_shift;
    s.m = m;
    s.a_mode = a_mode;
    s.cnt_lo = cnt_lo;
    s.cnt_hi = cnt_hi;
    s.d_value = d_value;
    s.seed = seed;
    KktProblem P;
    try {
      P = generate_kkt(s);
    } catch (const std::invalid_argument& e) {
      throw Error(kErrInvalid, e.what());
    }
    if (n_out) *n_out = P.K.n;
    if (nnz_out) *nnz_out = P.K.nnz();
    if (n1_out) *n1_out = P.n1;
    if (Ap) std::memcpy(Ap, P.K.p.data(), sizeof(int32_t) * P.K.p.size());
    if (Ai) std::memcpy(Ai, P.K.i.data(), sizeof(int32_t) * P.K.i.size());
    if (Ax) std::memcpy(Ax, P.K.x.data(), sizeof(double) * P.K.x.size());
    if (x_true) std::memcpy(x_true, P.x_true.data(), sizeof(double) * P.x_true.size());
    if (b) std::memcpy(b, P.b.data(), sizeof(double) * P.b.size());
    return 0;
  });
}

int mpir_amd_order(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* perm) {
  return guarded([&] {
    MPIR_CHECK(n >= 1 && Ap && Ai && perm, "bad ordering arguments");
    CscUpper A;
    A.n = n;
    A.p.assign(Ap, Ap + n + 1);
    A.i.assign(Ai, Ai + Ap[n]);
    A.x.assign(Ap[n], 0.0);
    std::vector<int32_t> p = amd_order(A);
    std::memcpy(perm, p.data(), sizeof(int32_t) * n);
    return 0;
  });
}

}  // extern "C"
