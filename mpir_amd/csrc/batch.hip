// Batches of independent quasi-definite systems on one device.
//
// A batch of S systems with up to R right-hand sides each is held as the
// block-diagonal system diag(K_1, ..., K_S) of N = sum n_s rows:
//   * the symbolic phase of every pattern comes from a PatternCache and the
//     plans are joined into one plan over the forest of elimination trees
//     (batch_plan.cpp), so the numeric factorization of every system runs in
//     one launch per level and step, and each application of M^{-1} in one
//     dependency-driven launch per direction and right-hand side;
//   * vectors are R slots of N entries (slot r holds right-hand side r of
//     every system; unused slots are zero);
//   * GMRES-IR runs one instance per (system, right-hand side) in lock step:
//     the SpMV, M^{-1}, the CGS2 products and the norms of all instances go
//     in shared launches, and every instance keeps its own scalars
//     (Hessenberg column, Givens rotations, stopping rules of
//     gmres_ir.hpp:534-741 in yex33/MPIR). An instance that has stopped is
//     masked: its vectors are no longer updated.
// Reductions are per instance over fixed chunks of its system's rows, added
// in chunk order, so an instance's arithmetic depends on its own system
// only: a system gives bit-identical results alone and in any batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

#include "kernels/norms.hpp"
#include "solver.hpp"

namespace mpir {

namespace {

constexpr int kChunk = 2048;  // rows per reduction chunk
constexpr int kBT = 256;      // threads per block of the chunk kernels
constexpr int kPerThread = kChunk / kBT;

// Device view of the batch layout.
struct Seg {
  int N, R, S, nch;
  const int32_t* sys_of_row;  // N
  const int32_t* ch_start;    // nch: first row of the chunk
  const int32_t* ch_len;      // nch
  const int32_t* sys_ch;      // S+1: chunks of each system
};

__device__ __forceinline__ double bsum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBT / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// NaN-propagating maximum over the block (kernels/norms.hpp)
__device__ __forceinline__ double bmax(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_down(v, off));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBT / 64; ++i) r = nan_max(r, sh[i]);
  __syncthreads();
  return r;
}

// Reduction over chunk ch of slot r. op 0: sum x y and op 2: max |x|, in
// part[r * nch + ch]; op 1: the 2-norm as a scaled sum of squares, the pair
// (scale, ssq) in part[2 (r * nch + ch)] and the entry after it.
__global__ void __launch_bounds__(kBT) k_bred_partial(Seg g, const double* __restrict__ x,
                                                      const double* __restrict__ y, int op, double* __restrict__ part) {
  __shared__ double sh[2 * kBT / 64];
  const int ch = blockIdx.x, r = blockIdx.y;
  const int64_t base = static_cast<int64_t>(r) * g.N + g.ch_start[ch];
  const int len = g.ch_len[ch];
  const int64_t slot = static_cast<int64_t>(r) * g.nch + ch;
  if (op == 1) {
    Ssq acc = ssq_init();
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int i = threadIdx.x + u * kBT;
      acc = ssq_add(acc, i < len ? x[base + i] : 0.0);
    }
    const Ssq t = block_ssq(acc, sh);
    if (threadIdx.x == 0) {
      part[2 * slot] = t.scale;
      part[2 * slot + 1] = t.ssq;
    }
    return;
  }
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int i = threadIdx.x + u * kBT;
    const double v = i < len ? x[base + i] : 0.0;
    if (op == 0) acc += v * (i < len ? y[base + i] : 0.0);
    else acc = nan_max(acc, fabs(v));
  }
  const double t = op == 2 ? bmax(acc, sh) : bsum(acc, sh);
  if (threadIdx.x == 0) part[slot] = t;
}

// out[r * S + s]: the chunks of system s in order (op 1: the 2-norm)
__global__ void k_bred_final(Seg g, const double* __restrict__ part, int op, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.R * g.S) return;
  const int r = q / g.S, s = q % g.S;
  const int64_t p0 = static_cast<int64_t>(r) * g.nch;
  if (op == 1) {
    Ssq acc = ssq_init();
    for (int ch = g.sys_ch[s]; ch < g.sys_ch[s + 1]; ++ch)
      acc = ssq_merge(acc, Ssq{part[2 * (p0 + ch)], part[2 * (p0 + ch) + 1]});
    out[q] = ssq_value(acc);
    return;
  }
  double acc = 0.0;
  for (int ch = g.sys_ch[s]; ch < g.sys_ch[s + 1]; ++ch) acc = op == 2 ? nan_max(acc, part[p0 + ch]) : acc + part[p0 + ch];
  out[q] = acc;
}

// Partial dots of w with basis vectors 0..k1-1 (stride ldv):
// part[(j * R + r) * nch + ch]
__global__ void __launch_bounds__(kBT) k_bmdot_partial(Seg g, const double* __restrict__ V, int64_t ldv, int k1,
                                                       const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[kBT / 64];
  const int ch = blockIdx.x, r = blockIdx.y;
  const int64_t base = static_cast<int64_t>(r) * g.N + g.ch_start[ch];
  const int len = g.ch_len[ch];
  double wv[kPerThread];
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int i = threadIdx.x + u * kBT;
    wv[u] = i < len ? w[base + i] : 0.0;
  }
  for (int j = 0; j < k1; ++j) {
    const double* vj = V + ldv * j + base;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int i = threadIdx.x + u * kBT;
      acc += (i < len ? vj[i] : 0.0) * wv[u];
    }
    const double t = bsum(acc, sh);
    if (threadIdx.x == 0) part[(static_cast<int64_t>(j) * g.R + r) * g.nch + ch] = t;
  }
}

// h[j * I + q] (+)= chunk sums of instance q = r * S + s, I = R * S
__global__ void k_bmdot_final(Seg g, const double* __restrict__ part, int k1, double* __restrict__ h, int accumulate) {
  const int I = g.R * g.S;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k1 * I) return;
  const int j = t / I, q = t % I, r = q / g.S, s = q % g.S;
  const double* p = part + (static_cast<int64_t>(j) * g.R + r) * g.nch;
  double acc = 0.0;
  for (int ch = g.sys_ch[s]; ch < g.sys_ch[s + 1]; ++ch) acc += p[ch];
  h[t] = accumulate ? h[t] + acc : acc;
}

__device__ __forceinline__ int instance_of(const Seg& g, int64_t t, int64_t* i_out) {
  const int r = static_cast<int>(t / g.N);
  const int64_t i = t - static_cast<int64_t>(r) * g.N;
  *i_out = i;
  return r * g.S + g.sys_of_row[i];
}

// w -= sum_j V_j h_j with the instance's h
__global__ void k_bsub_basis(Seg g, const double* __restrict__ V, int64_t ldv, int k1, const double* __restrict__ h,
                             double* __restrict__ w) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  int64_t i;
  const int q = instance_of(g, t, &i);
  const int I = g.R * g.S;
  double acc = 0.0;
  for (int j = 0; j < k1; ++j) acc += V[ldv * j + t] * h[static_cast<int64_t>(j) * I + q];
  w[t] -= acc;
}

// out = x * f[q]
__global__ void k_bscale(Seg g, const double* __restrict__ x, const double* __restrict__ f, double* __restrict__ out) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  int64_t i;
  const int q = instance_of(g, t, &i);
  out[t] = x[t] * f[q];
}

// x += d for the instances with mask[q]
__global__ void k_badd(Seg g, double* __restrict__ x, const double* __restrict__ d, const int32_t* __restrict__ mask) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  int64_t i;
  const int q = instance_of(g, t, &i);
  if (mask[q]) x[t] = x[t] + d[t];
}

// res = sum_{j < kq[q]} V_j y[j * I + q], j ascending (gmres_ir.hpp:735-739)
__global__ void k_bcombine(Seg g, const double* __restrict__ V, int64_t ldv, const int32_t* __restrict__ kq,
                           const double* __restrict__ y, double* __restrict__ res) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  int64_t i;
  const int q = instance_of(g, t, &i);
  const int I = g.R * g.S;
  double acc = 0.0;
  for (int j = 0; j < kq[q]; ++j) acc += V[ldv * j + t] * y[static_cast<int64_t>(j) * I + q];
  res[t] = acc;
}

// Slot r of combined row k (system s, local permuted row k - col_base[s]):
// b_s = (b o AD) in the fill-reducing order; zero for r >= nrhs[s].
// b of system s, right-hand side r: src[off[s] + r * n_s + lperm[k]].
__global__ void k_bperm_in(Seg g, const double* __restrict__ src, const int64_t* __restrict__ off,
                           const int32_t* __restrict__ nrhs, const int32_t* __restrict__ col_base,
                           const int32_t* __restrict__ lperm, const double* __restrict__ ad, double* __restrict__ out) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  const int r = static_cast<int>(t / g.N);
  const int64_t k = t - static_cast<int64_t>(r) * g.N;
  const int s = g.sys_of_row[k];
  const int64_t ns = col_base[s + 1] - col_base[s];
  const int64_t o = col_base[s] + lperm[k];  // original combined row
  out[t] = r < nrhs[s] ? src[off[s] + r * ns + lperm[k]] * ad[o] : 0.0;
}

// x of system s, right-hand side r < nrhs[s]: dst[off[s] + r n_s + lperm[k]] = xp * AD
__global__ void k_bperm_out(Seg g, const double* __restrict__ xp, const int64_t* __restrict__ off,
                            const int32_t* __restrict__ nrhs, const int32_t* __restrict__ col_base,
                            const int32_t* __restrict__ lperm, const double* __restrict__ ad, double* __restrict__ dst) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  const int r = static_cast<int>(t / g.N);
  const int64_t k = t - static_cast<int64_t>(r) * g.N;
  const int s = g.sys_of_row[k];
  if (r >= nrhs[s]) return;
  const int64_t ns = col_base[s + 1] - col_base[s];
  const int64_t o = col_base[s] + lperm[k];
  dst[off[s] + r * ns + lperm[k]] = xp[t] * ad[o];
}

// positive and zero pivots of every system: out[2 s], out[2 s + 1]
template <class T>
__global__ void __launch_bounds__(kBT) k_bpivots(const T* __restrict__ D, const int32_t* __restrict__ col_base,
                                                 int32_t* __restrict__ out) {
  __shared__ double sh[kBT / 64];
  const int s = blockIdx.x;
  double pos = 0.0, zero = 0.0;
  for (int i = col_base[s] + threadIdx.x; i < col_base[s + 1]; i += kBT) {
    const T d = D[i];
    pos += d > T(0) ? 1.0 : 0.0;
    zero += d == T(0) ? 1.0 : 0.0;
  }
  const double p = bsum(pos, sh), z = bsum(zero, sh);
  if (threadIdx.x == 0) {
    out[2 * s] = static_cast<int32_t>(p);
    out[2 * s + 1] = static_cast<int32_t>(z);
  }
}

inline unsigned nblk(int64_t n, int bs = 256) { return static_cast<unsigned>((n + bs - 1) / bs); }

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
void upload_vec(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (!v.empty()) MPIR_HIP(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace

int32_t GmresLdl::batch_size() const { return batch_ ? batch_->nsys : 0; }

int32_t GmresLdl::batch_n(int32_t s) const {
  MPIR_CHECK(batch_ && s >= 0 && s < batch_->nsys, "system index out of range");
  return batch_->col_base[s + 1] - batch_->col_base[s];
}

double* GmresLdl::bpinned(size_t count) {
  if (count > b_pinned_count_) {
    MPIR_HIP(hipStreamSynchronize(stream_));  // no copy may still use the old buffer
    if (b_pinned_) MPIR_HIP(hipHostFree(b_pinned_));
    b_pinned_ = nullptr;
    b_pinned_count_ = 0;
    const size_t c = std::max<size_t>(count, 4096);
    MPIR_HIP(hipHostMalloc(reinterpret_cast<void**>(&b_pinned_), c * sizeof(double), hipHostMallocDefault));
    b_pinned_count_ = c;
  }
  return b_pinned_;
}

// Host values to device memory through the pinned buffer. Waits for the
// stream first, so no earlier copy still reads the buffer.
void GmresLdl::put_host(const void* src, size_t bytes, void* dst) {
  double* hp = bpinned((bytes + sizeof(double) - 1) / sizeof(double));
  MPIR_HIP(hipStreamSynchronize(stream_));
  std::memcpy(hp, src, bytes);
  MPIR_HIP(hipMemcpyAsync(dst, hp, bytes, hipMemcpyHostToDevice, stream_));
}

void GmresLdl::analyze_batch(const std::vector<std::shared_ptr<const PatternEntry>>& parts) {
  const double t0 = now_ms();
  MPIR_CHECK(!parts.empty(), "a batch needs at least one system");
  MPIR_CHECK(precond_ == kPrecLdl, "batches use the L D L^T preconditioner");
  MPIR_CHECK(uw_ == kF64 && ur_ == kF64, "batches refine in f64 working and residual precision");
  std::vector<const PatternEntry*> raw;
  for (const auto& p : parts) {
    MPIR_CHECK(p != nullptr, "null pattern");
    raw.push_back(p.get());
  }
  analyzed_ = factored_ = false;
  precond_only_ = false;
  // the joined system and plan reuse the memory of the previous batch
  if (!batch_) batch_ = std::make_unique<BatchSystem>();
  if (!plan_) plan_ = std::make_unique<FactorPlan>();
  BatchSystem* B = batch_.get();
  try {
    combine_patterns(raw, *B, *plan_);
    apply_tuning();
  } catch (const std::exception& e) {
    batch_.reset();
    throw Error(kErrInvalid, e.what());
  }
  batch_parts_ = parts;
  hipStream_t st = stream_;
  n_ = B->col_base.back();
  nnz_ = B->ent_base.back();
  Ap_.clear();
  Ai_.clear();
  Cp_.clear();
  Ci_.clear();
  parent_.clear();
  Lnz_.clear();
  perm_ = B->perm;
  relaxed_ = B->relaxed;
  true_nnz_l_ = B->nnz_l;
  true_flops_ = B->flops;
  upload_plan();
  upload_vec(d_csr_p_, B->csr_p, st);
  upload_vec(d_csr_j_, B->csr_j, st);
  upload_vec(d_csr_src_, B->csr_src, st);
  d_csr_x_.alloc(std::max<size_t>(1, B->csr_j.size()) * sizeof(double));
  upload_vec(d_entry_col_, B->ecol, st);
  upload_vec(d_entry_row_, B->erow, st);
  upload_vec(d_diag_pos_, B->diag_pos, st);
  upload_vec(d_perm_, perm_, st);
  // reduction chunks: every system's rows cut into pieces of kChunk
  std::vector<int32_t> ch_start, ch_len, sys_ch(1, 0);
  for (int32_t s = 0; s < B->nsys; ++s) {
    for (int32_t r = B->col_base[s]; r < B->col_base[s + 1]; r += kChunk) {
      ch_start.push_back(r);
      ch_len.push_back(std::min(kChunk, B->col_base[s + 1] - r));
    }
    sys_ch.push_back(static_cast<int32_t>(ch_start.size()));
  }
  b_nch_ = static_cast<int32_t>(ch_start.size());
  upload_vec(d_b_sys_of_row_, B->sys_of_row, st);
  upload_vec(d_b_col_base_, B->col_base, st);
  upload_vec(d_b_lperm_, B->local_perm, st);
  upload_vec(d_b_ch_start_, ch_start, st);
  upload_vec(d_b_ch_len_, ch_len, st);
  upload_vec(d_b_sys_ch_, sys_ch, st);
  d_b_piv_.alloc(2 * sizeof(int32_t) * B->nsys);
  alloc_numeric();
  MPIR_HIP(hipStreamSynchronize(st));

  const Supernodes& S = plan_->sn;
  stats_ = SolveStats();
  stats_.n = n_;
  stats_.nnz = nnz_;
  stats_.nnz_l = true_nnz_l_;
  stats_.nsupernodes = S.ns;
  stats_.nlevels = S.nlevels;
  stats_.max_front = S.max_front;
  stats_.factor_flops = true_flops_;
  stats_.device_bytes = static_cast<int64_t>(d_panel_.bytes() + d_upool_.bytes());
  bstats_ = BatchStats();
  bstats_.nsys = batch_->nsys;
  analyzed_ = true;
  stats_.analyze_ms = now_ms() - t0;
}

void GmresLdl::factor_batch(const double* const* Ax, bool device_ptr) {
  if (!analyzed_ || !batch_) throw Error(kErrState, "factor_batch called before analyze_batch");
  MPIR_CHECK(Ax != nullptr, "null value array list");
  const double t0 = now_ms();
  factored_ = false;
  for (int32_t s = 0; s < batch_->nsys; ++s) {
    const int64_t cnt = batch_->ent_base[s + 1] - batch_->ent_base[s];
    MPIR_CHECK(Ax[s] != nullptr, "null value array");
    MPIR_HIP(hipMemcpyAsync(d_ax_.as<double>() + batch_->ent_base[s], Ax[s], cnt * sizeof(double),
                            device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  }
  factor_loaded(t0);
  bstats_.factor_ms = stats_.factor_ms;
}

void GmresLdl::batch_count_pivots() {
  const int S = batch_->nsys;
  if (uf_ == kF64)
    hipLaunchKernelGGL(k_bpivots<double>, dim3(S), dim3(kBT), 0, stream_, d_D_.as<double>(), d_b_col_base_.as<int32_t>(),
                       d_b_piv_.as<int32_t>());
  else
    hipLaunchKernelGGL(k_bpivots<float>, dim3(S), dim3(kBT), 0, stream_, d_D_.as<float>(), d_b_col_base_.as<int32_t>(),
                       d_b_piv_.as<int32_t>());
  int32_t* hp = reinterpret_cast<int32_t*>(bpinned((2 * S + 1) / 2 + 1));
  MPIR_HIP(hipMemcpyAsync(hp, d_b_piv_.get(), 2 * sizeof(int32_t) * S, hipMemcpyDeviceToHost, stream_));
}

void GmresLdl::batch_collect_pivots(bool zero_pivot) {
  const int S = batch_->nsys;
  const int32_t* hp = reinterpret_cast<const int32_t*>(b_pinned_);
  bstats_.positive_pivots.assign(S, 0);
  bstats_.zero_pivots.assign(S, 0);
  int first_bad = -1;
  for (int s = 0; s < S; ++s) {
    bstats_.positive_pivots[s] = hp[2 * s];
    bstats_.zero_pivots[s] = hp[2 * s + 1];
    if (hp[2 * s + 1] > 0 && first_bad < 0) first_bad = s;
  }
  if (zero_pivot && first_bad >= 0)
    throw Error(kErrNumeric, "zero pivot in the LDL^T factorization of system " + std::to_string(first_bad));
}

// M^{-1} applied to R slots of n entries: each tile of the factor is read
// once for up to 8 right-hand sides
void GmresLdl::precond_apply_slots(double* x, int R) {
  if (precond_ == kPrecIlu) {
    for (int r = 0; r < R; ++r) precond_apply_dev(x + static_cast<int64_t>(r) * n_);
    return;
  }
  set_solve_slots(std::min(R, 8));
  stats_.precond_applications += R;
  double* u = d_uvec_.as<double>();
  if (uf_ == kF64)
    ldl_solve_multi<double, double, double>(solve_args<double, double>(d_panel_.get()), dsp_, x, R, u, stream_);
  else if (uf_ == kF32)
    ldl_solve_multi<float, float, double>(solve_args<float, float>(d_panel_.get()), dsp_, x, R, u, stream_);
  else if (uf_ == kF16)
    ldl_solve_multi<__half, float, double>(solve_args<__half, float>(d_panel16_.get()), dsp_, x, R, u, stream_);
  else
    ldl_solve_multi<__hip_bfloat16, float, double>(solve_args<__hip_bfloat16, float>(d_panel16_.get()), dsp_, x, R, u,
                                                   stream_);
  bstats_.precond_sweeps++;
}

void GmresLdl::spmv_slots(const double* x, double* y, int R) {
  for (int r = 0; r < R; ++r) spmv_dev(x + static_cast<int64_t>(r) * n_, y + static_cast<int64_t>(r) * n_);
}

void GmresLdl::residual_slots(const double* x, const double* b, double* r, int R) {
  CsrDev A{n_, d_csr_p_.as<int32_t>(), d_csr_j_.as<int32_t>(), d_csr_x_.as<double>()};
  for (int k = 0; k < R; ++k) {
    const int64_t o = static_cast<int64_t>(k) * n_;
    residual_f64(A, x + o, b + o, r + o, stream_);
  }
}

// per-instance norms (op 1: 2-norm, 2: inf-norm) of R slots, copied to the host
void GmresLdl::bnorms(const double* v, int op, int R, std::vector<double>& out) {
  const int S = batch_->nsys;
  const int I = R * S;
  Seg g{n_, R, S, b_nch_, d_b_sys_of_row_.as<int32_t>(), d_b_ch_start_.as<int32_t>(), d_b_ch_len_.as<int32_t>(),
        d_b_sys_ch_.as<int32_t>()};
  d_red_.ensure(static_cast<size_t>(2) * R * b_nch_ * sizeof(double));
  double* part = d_red_.as<double>();
  hipLaunchKernelGGL(k_bred_partial, dim3(b_nch_, R), dim3(kBT), 0, stream_, g, v, nullptr, op, part);
  hipLaunchKernelGGL(k_bred_final, dim3(nblk(I)), dim3(256), 0, stream_, g, part, op, d_b_scalar_.as<double>());
  double* hp = bpinned(I);
  MPIR_HIP(hipMemcpyAsync(hp, d_b_scalar_.get(), I * sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPIR_HIP(hipStreamSynchronize(stream_));
  out.assign(hp, hp + I);
}

// PrecondGmres (gmres_ir.hpp:636-741) for every active instance at once:
// res = V y with V the CGS2 Arnoldi basis of M^{-1} K_s from M^{-1} b.
void GmresLdl::precond_gmres_batch(const double* b, double* res, const std::vector<char>& active) {
  hipStream_t st = stream_;
  const int S = batch_->nsys;
  const int R = b_slots_;
  const int I = R * S;
  const int64_t N = n_;
  const int64_t RN = static_cast<int64_t>(R) * N;
  // the reference caps the cycle at n (gmres_ir.hpp:643), per system here
  std::vector<int> m_of(I);
  int M = 0;
  for (int q = 0; q < I; ++q) {
    const int s = q % S;
    m_of[q] = static_cast<int>(std::min<int64_t>(max_gmres_, batch_n(s)));
    if (active[q]) M = std::max(M, m_of[q]);
  }
  MPIR_CHECK(M + 2 <= 4096, "GMRES iteration limit above 4094");
  const int64_t ldv = RN;
  d_V_.ensure(static_cast<size_t>(M + 1) * ldv * sizeof(double));
  d_hcol_.ensure(static_cast<size_t>(M + 2) * I * sizeof(double));
  d_small_.ensure(static_cast<size_t>(M + 2) * I * sizeof(double));
  d_red_.ensure(static_cast<size_t>(std::max(M + 1, 2)) * R * b_nch_ * sizeof(double));
  d_b_scalar_.ensure(static_cast<size_t>(std::max(M, 1)) * I * sizeof(double));
  d_b_kq_.ensure(I * sizeof(int32_t));
  double* V = d_V_.as<double>();
  double* hc = d_hcol_.as<double>();
  double* h2 = d_small_.as<double>();
  double* part = d_red_.as<double>();
  Seg g{n_, R, S, b_nch_, d_b_sys_of_row_.as<int32_t>(), d_b_ch_start_.as<int32_t>(), d_b_ch_len_.as<int32_t>(),
        d_b_sys_ch_.as<int32_t>()};
  MPIR_HIP(hipMemsetAsync(res, 0, RN * sizeof(double), st));
  std::vector<double> rho;
  bnorms(b, 1, R, rho);
  std::vector<char> on(I, 0);
  bool any = false;
  for (int q = 0; q < I; ++q) {
    on[q] = active[q] && !(rho[q] < tol_);
    any = any || on[q];
  }
  if (!any) return;
  MPIR_HIP(hipMemcpyAsync(V, b, RN * sizeof(double), hipMemcpyDeviceToDevice, st));
  precond_apply_slots(V, R);
  bnorms(V, 1, R, rho);
  // V_0 = M^{-1} b / rho (instances not iterating: zero)
  std::vector<double> fac(I, 0.0);
  for (int q = 0; q < I; ++q) fac[q] = on[q] ? 1.0 / rho[q] : 0.0;
  put_host(fac.data(), I * sizeof(double), d_b_scalar_.get());
  hipLaunchKernelGGL(k_bscale, dim3(nblk(RN)), dim3(256), 0, st, g, V, d_b_scalar_.as<double>(), V);
  // per instance: Hessenberg (column-major (M+1) x M), Givens, g
  std::vector<double> H(static_cast<size_t>(I) * (M + 1) * std::max(M, 1), 0.0);
  std::vector<double> gv(static_cast<size_t>(I) * (M + 1), 0.0), cs(static_cast<size_t>(I) * M, 0.0),
      sn(static_cast<size_t>(I) * M, 0.0);
  std::vector<int> kq(I, 0);
  auto Hq = [&](int q, int i, int j) -> double& { return H[(static_cast<size_t>(q) * std::max(M, 1) + j) * (M + 1) + i]; };
  for (int q = 0; q < I; ++q) {
    gv[static_cast<size_t>(q) * (M + 1)] = rho[q];
  }
  std::vector<char> run = on;  // still iterating
  for (int k = 0; k < M; ++k) {
    bool more = false;
    for (int q = 0; q < I; ++q) {
      run[q] = run[q] && k < m_of[q];
      more = more || run[q];
    }
    if (!more) break;
    bstats_.arnoldi_steps++;
    double* w = V + ldv * (k + 1);
    spmv_slots(V + ldv * k, w, R);
    precond_apply_slots(w, R);
    const int k1 = k + 1;
    for (int pass = 0; pass < 2; ++pass) {
      double* hp = pass == 0 ? hc : h2;
      hipLaunchKernelGGL(k_bmdot_partial, dim3(b_nch_, R), dim3(kBT), 0, st, g, V, ldv, k1, w, part);
      hipLaunchKernelGGL(k_bmdot_final, dim3(nblk(static_cast<int64_t>(k1) * I)), dim3(256), 0, st, g, part, k1, hp, 0);
      hipLaunchKernelGGL(k_bsub_basis, dim3(nblk(RN)), dim3(256), 0, st, g, V, ldv, k1, hp, w);
    }
    // the column is h = h1 + h2, summed on the host below
    hipLaunchKernelGGL(k_bred_partial, dim3(b_nch_, R), dim3(kBT), 0, st, g, w, nullptr, 1, part);
    hipLaunchKernelGGL(k_bred_final, dim3(nblk(I)), dim3(256), 0, st, g, part, 1, hc + static_cast<int64_t>(k1) * I);
    // both passes' coefficients and the norm to the host
    double* hs = bpinned(static_cast<size_t>(2 * k1 + 1) * I);
    MPIR_HIP(hipMemcpyAsync(hs, hc, static_cast<size_t>(k1 + 1) * I * sizeof(double), hipMemcpyDeviceToHost, st));
    double* hs2 = hs + static_cast<size_t>(k1 + 1) * I;
    MPIR_HIP(hipMemcpyAsync(hs2, h2, static_cast<size_t>(k1) * I * sizeof(double), hipMemcpyDeviceToHost, st));
    MPIR_HIP(hipStreamSynchronize(st));
    std::fill(fac.begin(), fac.end(), 0.0);
    for (int q = 0; q < I; ++q) {
      if (!run[q]) continue;
      double* gq = &gv[static_cast<size_t>(q) * (M + 1)];
      double* cq = &cs[static_cast<size_t>(q) * M];
      double* sq = &sn[static_cast<size_t>(q) * M];
      for (int j = 0; j <= k; ++j) Hq(q, j, k) = hs[static_cast<size_t>(j) * I + q] + hs2[static_cast<size_t>(j) * I + q];
      Hq(q, k + 1, k) = hs[static_cast<size_t>(k1) * I + q];
      fac[q] = Hq(q, k + 1, k) != 0 ? 1.0 / Hq(q, k + 1, k) : 1.0;
      for (int j = 0; j < k; ++j) {
        const double w1 = cq[j] * Hq(q, j, k) - sq[j] * Hq(q, j + 1, k);
        const double w2 = sq[j] * Hq(q, j, k) + cq[j] * Hq(q, j + 1, k);
        Hq(q, j, k) = w1;
        Hq(q, j + 1, k) = w2;
      }
      const double nu = dnrm2_pair(Hq(q, k, k), Hq(q, k + 1, k));
      if (nu != 0) {
        cq[k] = Hq(q, k, k) / nu;
        sq[k] = -Hq(q, k + 1, k) / nu;
        Hq(q, k, k) = cq[k] * Hq(q, k, k) - sq[k] * Hq(q, k + 1, k);
        Hq(q, k + 1, k) = 0.0;
        const double w1 = cq[k] * gq[k] - sq[k] * gq[k + 1];
        const double w2 = sq[k] * gq[k] + cq[k] * gq[k + 1];
        gq[k] = w1;
        gq[k + 1] = w2;
      }
      kq[q] = k + 1;
      run[q] = std::fabs(gq[k + 1]) > tol_;
    }
    put_host(fac.data(), I * sizeof(double), d_b_scalar_.get());
    hipLaunchKernelGGL(k_bscale, dim3(nblk(RN)), dim3(256), 0, st, g, w, d_b_scalar_.as<double>(), w);
  }
  // back substitution per instance, then res = V y
  int kmax = 0;
  std::vector<double> y(static_cast<size_t>(std::max(M, 1)) * I, 0.0);
  for (int q = 0; q < I; ++q) {
    const int k = kq[q];
    kmax = std::max(kmax, k);
    bstats_.gmres_iterations[q] += k;
    const double* gq = &gv[static_cast<size_t>(q) * (M + 1)];
    std::vector<double> yq(std::max(k, 1), 0.0);
    for (int i = k; i-- > 0;) {
      yq[i] = gq[i];
      for (int j = i + 1; j < k; ++j) yq[i] -= Hq(q, i, j) * yq[j];
      yq[i] /= Hq(q, i, i);
    }
    for (int j = 0; j < k; ++j) y[static_cast<size_t>(j) * I + q] = yq[j];
  }
  if (kmax == 0) {
    MPIR_HIP(hipStreamSynchronize(st));
    return;
  }
  // y (kmax x I) and the step counts in one upload
  std::vector<double> yk(static_cast<size_t>(kmax) * I + (I + 1) / 2, 0.0);
  std::memcpy(yk.data(), y.data(), static_cast<size_t>(kmax) * I * sizeof(double));
  int32_t* hk = reinterpret_cast<int32_t*>(yk.data() + static_cast<size_t>(kmax) * I);
  for (int q = 0; q < I; ++q) hk[q] = kq[q];
  d_b_scalar2_.ensure(yk.size() * sizeof(double));
  put_host(yk.data(), yk.size() * sizeof(double), d_b_scalar2_.get());
  const double* dy = d_b_scalar2_.as<double>();
  const int32_t* dk = reinterpret_cast<const int32_t*>(dy + static_cast<size_t>(kmax) * I);
  hipLaunchKernelGGL(k_bcombine, dim3(nblk(RN)), dim3(256), 0, st, g, V, ldv, dk, dy, res);
  MPIR_HIP(hipStreamSynchronize(st));
}

// GmresLDLIR::Solve (gmres_ir.hpp:534-634) for every (system, right-hand
// side) instance of the batch in lock step.
void GmresLdl::solve_batch(const int32_t* nrhs, const double* const* b, double* const* x, bool device_ptr) {
  if (!factored_ || !batch_) throw Error(kErrState, "solve_batch called before a successful factor_batch");
  MPIR_CHECK(nrhs && b && x, "null batch arguments");
  const double t_start = now_ms();
  hipStream_t st = stream_;
  const int S = batch_->nsys;
  int R = 0;
  for (int s = 0; s < S; ++s) {
    MPIR_CHECK(nrhs[s] >= 0, "negative right-hand-side count");
    MPIR_CHECK(nrhs[s] == 0 || (b[s] && x[s]), "null right-hand side or solution");
    R = std::max(R, nrhs[s]);
  }
  MPIR_CHECK(R >= 1, "no right-hand sides in the batch");
  b_slots_ = R;
  const int I = R * S;
  const int64_t N = n_;
  const int64_t RN = static_cast<int64_t>(R) * N;
  // per-system right-hand sides packed into one device array (ld = n_s)
  std::vector<int64_t> off(S + 1, 0);
  for (int s = 0; s < S; ++s) off[s + 1] = off[s] + static_cast<int64_t>(nrhs[s]) * batch_n(s);
  d_b_vec_.ensure(static_cast<size_t>(5 * RN + 2 * off[S]) * sizeof(double));
  double* vb = d_b_vec_.as<double>();
  double* vx = vb + RN;
  double* vr = vx + RN;
  double* vd = vr + RN;
  double* vt = vd + RN;
  double* io_in = vt + RN;
  double* io_out = io_in + off[S];
  for (int s = 0; s < S; ++s)
    if (nrhs[s] > 0)
      MPIR_HIP(hipMemcpyAsync(io_in + off[s], b[s], (off[s + 1] - off[s]) * sizeof(double),
                              device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  std::vector<int32_t> nr(nrhs, nrhs + S);
  upload_vec(d_b_off_, off, st);
  upload_vec(d_b_nrhs_, nr, st);
  d_b_mask_.ensure(I * sizeof(int32_t));
  d_b_scalar_.ensure(I * sizeof(double));
  Seg g{n_, R, S, b_nch_, d_b_sys_of_row_.as<int32_t>(), d_b_ch_start_.as<int32_t>(), d_b_ch_len_.as<int32_t>(),
        d_b_sys_ch_.as<int32_t>()};
  d_red_.ensure(static_cast<size_t>(R) * b_nch_ * sizeof(double) * 2);
  hipLaunchKernelGGL(k_bperm_in, dim3(nblk(RN)), dim3(256), 0, st, g, io_in, d_b_off_.as<int64_t>(),
                     d_b_nrhs_.as<int32_t>(), d_b_col_base_.as<int32_t>(), d_b_lperm_.as<int32_t>(), d_ad_.as<double>(),
                     vb);
  // statistics (instances q = r S + s internally)
  bstats_.max_rhs = R;
  bstats_.arnoldi_steps = bstats_.precond_sweeps = 0;
  bstats_.gmres_iterations.assign(I, 0);
  std::vector<int32_t> ir_it(I, 0), stop(I, 0);
  std::vector<double> relres(I, 0.0), done(I, 0.0);
  std::vector<char> valid(I, 0);
  for (int q = 0; q < I; ++q) valid[q] = (q / S) < nrhs[q % S];

  precond_gmres_batch(vb, vx, valid);
  residual_slots(vx, vb, vr, R);
  std::vector<double> bnorm, rn, dn, xn;
  bnorms(vb, 2, R, bnorm);
  std::vector<double> r_prev(I, std::numeric_limits<double>::max());
  std::vector<char> act = valid;
  // an instance whose right-hand side or residual is not finite stops with
  // reason 4 (a single-system handle reports MPIR_ERR_NUMERIC instead; one
  // bad instance must not end the others of its batch)
  for (int q = 0; q < I; ++q)
    if (act[q] && !std::isfinite(bnorm[q])) {
      stop[q] = 4;
      relres[q] = std::numeric_limits<double>::quiet_NaN();
      act[q] = 0;
    }
  int64_t iter;
  for (iter = 1; iter <= max_ir_; ++iter) {
    bool any = false;
    for (int q = 0; q < I; ++q) any = any || act[q];
    if (!any) break;
    std::vector<double> scale, f(I);
    bnorms(vr, 2, R, scale);
    for (int q = 0; q < I; ++q) {
      if (act[q] && scale[q] == 0.0) {
        // an exactly zero residual: x is exact (the reference would scale by 1/0)
        stop[q] = 3;
        ir_it[q] = static_cast<int32_t>(iter);
        relres[q] = 0.0;
        act[q] = 0;
        done[q] = now_ms() - t_start;
      }
      f[q] = act[q] ? 1.0 / scale[q] : 0.0;
    }
    put_host(f.data(), I * sizeof(double), d_b_scalar_.get());
    hipLaunchKernelGGL(k_bscale, dim3(nblk(RN)), dim3(256), 0, st, g, vr, d_b_scalar_.as<double>(), vt);
    precond_gmres_batch(vt, vd, act);
    for (int q = 0; q < I; ++q) f[q] = act[q] ? scale[q] : 0.0;
    put_host(f.data(), I * sizeof(double), d_b_scalar_.get());
    hipLaunchKernelGGL(k_bscale, dim3(nblk(RN)), dim3(256), 0, st, g, vd, d_b_scalar_.as<double>(), vd);
    bnorms(vd, 2, R, dn);
    bnorms(vx, 2, R, xn);
    std::vector<int32_t> hm(I);
    for (int q = 0; q < I; ++q) hm[q] = act[q] ? 1 : 0;
    put_host(hm.data(), I * sizeof(int32_t), d_b_mask_.get());
    hipLaunchKernelGGL(k_badd, dim3(nblk(RN)), dim3(256), 0, st, g, vx, vd, d_b_mask_.as<int32_t>());
    residual_slots(vx, vb, vr, R);
    bnorms(vr, 2, R, rn);
    for (int q = 0; q < I; ++q) {
      if (!act[q]) continue;
      const double r_norm = rn[q] / bnorm[q];
      ir_it[q] = static_cast<int32_t>(iter);
      relres[q] = r_norm;
      if (!std::isfinite(r_norm) || !std::isfinite(dn[q])) {
        stop[q] = 4;
      } else if (dn[q] <= xn[q] * (10 * std::numeric_limits<double>::epsilon())) {
        stop[q] = 1;
      } else if (std::fabs(r_norm - r_prev[q]) < tol_) {
        stop[q] = 2;
      } else {
        r_prev[q] = r_norm;
        continue;
      }
      act[q] = 0;
      done[q] = now_ms() - t_start;
    }
  }
  if (max_ir_ == 0) {
    bnorms(vr, 2, R, rn);
    for (int q = 0; q < I; ++q) relres[q] = rn[q] / bnorm[q];
  }
  const double t_end = now_ms() - t_start;
  for (int q = 0; q < I; ++q)
    if (valid[q] && act[q]) done[q] = t_end;  // iteration limit
  hipLaunchKernelGGL(k_bperm_out, dim3(nblk(RN)), dim3(256), 0, st, g, vx, d_b_off_.as<int64_t>(),
                     d_b_nrhs_.as<int32_t>(), d_b_col_base_.as<int32_t>(), d_b_lperm_.as<int32_t>(), d_ad_.as<double>(),
                     io_out);
  for (int s = 0; s < S; ++s)
    if (nrhs[s] > 0)
      MPIR_HIP(hipMemcpyAsync(x[s], io_out + off[s], (off[s + 1] - off[s]) * sizeof(double),
                              device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
  sync_check();
  // per-instance statistics in system-major order (s * R + r)
  bstats_.ir_iterations.assign(I, 0);
  bstats_.stop_reason.assign(I, 0);
  bstats_.final_rel_residual.assign(I, 0.0);
  bstats_.done_ms.assign(I, 0.0);
  std::vector<int32_t> gm(I, 0);
  for (int q = 0; q < I; ++q) {
    const int r = q / S, s = q % S, o = s * R + r;
    bstats_.ir_iterations[o] = ir_it[q];
    bstats_.stop_reason[o] = stop[q];
    bstats_.final_rel_residual[o] = relres[q];
    bstats_.done_ms[o] = done[q];
    gm[o] = bstats_.gmres_iterations[q];
  }
  bstats_.gmres_iterations.swap(gm);
  bstats_.solve_ms = now_ms() - t_start;
  stats_.solve_ms = bstats_.solve_ms;
}

}  // namespace mpir
