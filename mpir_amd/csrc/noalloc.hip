// M^{-1} applications that allocate nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).
//
// The triangular solves need work memory beyond their input and output:
// per-front completion counters and an error flag, partial sums of the
// big-front products, and the fronts' update vectors. GmresLdl::
// precond_apply_dev takes them from buffers the handle allocated; here they
// come from `scratch`, a file-scope device array of kScratchBytes that the
// code object carries, so the call itself performs no device allocation.
//
// Every GPU holds its own copy of the array, at its own address, from the
// first use on that device until the process ends. In host code `scratch`
// names a host-side stand-in, not a device address: the device address is
// looked up once per device with hipGetSymbolAddress (with that device
// current) and only that address is used afterwards. All calls on one
// device share the array and its contents persist between calls: the
// counters and the error flag are cleared by a kernel at the start of each
// application (a fill kernel, not hipMemsetAsync, so a captured graph stays
// valid), and everything else is written before it is read. Calls on one
// device are serialized by a lock held until their result is complete;
// overlapping use of the array from outside these calls would race.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <mutex>
#include <string>

#include "solver.hpp"

namespace mpir {

namespace {

constexpr int64_t kScratchBytes = int64_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_lookup_mu;
std::array<void*, kMaxDevices> g_scratch_ptr{};  // device address of scratch, per device
std::array<std::mutex, kMaxDevices> g_device_mu;  // one call at a time per device

// The device address of this device's scratch (looked up once per device).
void* device_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_lookup_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    MPIR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MPIR_CHECK(p != nullptr, "no device address for the static scratch");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

inline int64_t round256(int64_t b) { return (b + 255) / 256 * 256; }

}  // namespace

int64_t GmresLdl::precond_work_bytes() const {
  MPIR_CHECK(plan_ != nullptr, "no factorization plan");
  const FactorPlan& P = *plan_;
  const SolveDev& d = dsp_.d;
  const int64_t rows = std::max<int64_t>(1, static_cast<int64_t>(P.sn.rows.size()));
  const int64_t parts = P.nypart + P.nupart + P.nrpart + P.nxpart + 64;
  return round256((static_cast<int64_t>(d.ncounters) + 1) * static_cast<int64_t>(sizeof(int))) +
         round256(parts * static_cast<int64_t>(sizeof(double))) + round256(rows * static_cast<int64_t>(sizeof(double)));
}

void GmresLdl::precond_apply_in(double* x, void* work) {
  MPIR_CHECK(factored_ && precond_ == kPrecLdl && !batch_, "needs a single-system L D L^T factor");
  const FactorPlan& P = *plan_;
  SolveDev g = dsp_.d;  // the plan's task lists and offsets; work areas from `work`
  char* base = static_cast<char*>(work);
  auto take = [&](int64_t bytes) {
    char* p = base;
    base += round256(bytes);
    return p;
  };
  // counters, then the error flag: the solve's own fill kernel clears both
  g.counters = reinterpret_cast<int*>(take((static_cast<int64_t>(g.ncounters) + 1) * static_cast<int64_t>(sizeof(int))));
  g.err = g.counters + g.ncounters;
  g.ncounters += 1;
  double* wk = reinterpret_cast<double*>(take((P.nypart + P.nupart + P.nrpart + P.nxpart + 64) *
                                              static_cast<int64_t>(sizeof(double))));
  g.ypart = wk;
  g.upart = wk + P.nypart;
  g.rpart = wk + P.nypart + P.nupart;
  g.xpart = wk + P.nypart + P.nupart + P.nrpart;
  g.syp = P.nypart;
  g.sup = P.nupart;
  g.srp = P.nrpart;
  g.sxp = P.nxpart;
  const int64_t rows = std::max<int64_t>(1, static_cast<int64_t>(P.sn.rows.size()));
  double* u = reinterpret_cast<double*>(take(rows * static_cast<int64_t>(sizeof(double))));
  g.su = rows;
  g.sx = n_;
  stats_.precond_applications++;
  if (uf_ == kF64) ldl_solve_on<double, double, double>(solve_args<double, double>(d_panel_.get()), dsp_, g, x, u, stream_);
  else if (uf_ == kF32) ldl_solve_on<float, float, double>(solve_args<float, float>(d_panel_.get()), dsp_, g, x, u, stream_);
  else if (uf_ == kF16)
    ldl_solve_on<__half, float, double>(solve_args<__half, float>(d_panel16_.get()), dsp_, g, x, u, stream_);
  else
    ldl_solve_on<__hip_bfloat16, float, double>(solve_args<__hip_bfloat16, float>(d_panel16_.get()), dsp_, g, x, u,
                                                stream_);
  // the error flag through the handle's pinned word (no staging copy)
  MPIR_HIP(hipMemcpyAsync(&h_pinned_[3], g.err, sizeof(int), hipMemcpyDeviceToHost, stream_));
  MPIR_HIP(hipStreamSynchronize(stream_));
  int e = 0;
  std::memcpy(&e, &h_pinned_[3], sizeof(int));
  if (e) throw Error(kErrInternal, "a triangular-solve dependency wait expired");
}

void precond_apply_noalloc(GmresLdl& s, double* x) {
  MPIR_CHECK(x != nullptr, "null vector");
  int dev = 0;
  MPIR_HIP(hipGetDevice(&dev));
  MPIR_CHECK(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  const int64_t need = s.precond_work_bytes();
  if (need > kScratchBytes)
    throw Error(kErrInvalid, "this factor's triangular solves need " + std::to_string(need) +
                                 " bytes of work memory; the static scratch holds " + std::to_string(kScratchBytes) +
                                 " (use mpir_gmres_ldl_precond_dev)");
  void* p = device_scratch(dev);
  std::lock_guard<std::mutex> lk(g_device_mu[dev]);
  s.precond_apply_in(x, p);
}

}  // namespace mpir
