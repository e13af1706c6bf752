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

#include <atomic>
#include <cstring>
#include <thread>

#include <algorithm>
#include <cfloat>
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

// every kernel of a solve takes the flag of the work it belongs to (a
// refinement iteration, a GMRES cycle, an Arnoldi step) and returns at once
// when no instance takes part in it (flag zero); null: always runs
__device__ __forceinline__ bool gated(const int* gate) { return gate != nullptr && *gate == 0; }

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
                                                      const double* __restrict__ y, int op, double* __restrict__ part,
                                                      const int* __restrict__ gate) {
  __shared__ double sh[2 * kBT / 64];
  if (gated(gate)) return;
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
__global__ void k_bred_final(Seg g, const double* __restrict__ part, int op, double* __restrict__ out,
                             const int* __restrict__ gate) {
  if (gated(gate)) return;
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
                                                       const double* __restrict__ w, double* __restrict__ part,
                                                       const int* __restrict__ gate) {
  __shared__ double sh[kBT / 64];
  if (gated(gate)) return;
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
__global__ void k_bmdot_final(Seg g, const double* __restrict__ part, int k1, double* __restrict__ h, int accumulate,
                              const int* __restrict__ gate) {
  if (gated(gate)) return;
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
                             double* __restrict__ w, const int* __restrict__ gate) {
  if (gated(gate)) return;
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
__global__ void k_bscale(Seg g, const double* __restrict__ x, const double* __restrict__ f, double* __restrict__ out,
                         const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  int64_t i;
  const int q = instance_of(g, t, &i);
  out[t] = x[t] * f[q];
}

// x += d for the instances with mask[q]
__global__ void k_badd(Seg g, double* __restrict__ x, const double* __restrict__ d, const int32_t* __restrict__ mask,
                       const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(g.R) * g.N) return;
  int64_t i;
  const int q = instance_of(g, t, &i);
  if (mask[q]) x[t] = x[t] + d[t];
}

// res = sum_{j < kq[q]} V_j y[j * I + q], j ascending (gmres_ir.hpp:735-739)
__global__ void k_bcombine(Seg g, const double* __restrict__ V, int64_t ldv, const int32_t* __restrict__ kq,
                           const double* __restrict__ y, double* __restrict__ res, const int* __restrict__ gate) {
  if (gated(gate)) return;
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

// ---------------------------------------------------------------- device-resident GMRES-IR
// Every scalar of the batched GMRES-IR lives on the device, per instance q
// (I = R S): the refinement state (gmres_ir.hpp:534-634), the Hessenberg
// matrix, Givens rotations and g of the current GMRES cycle
// (gmres_ir.hpp:636-741). Flags say which refinement iterations, GMRES
// cycles and Arnoldi steps any instance takes part in; each kernel reads
// its flag first, so the whole solve is one fixed launch sequence up to the
// iteration limits, which is captured once as a HIP graph.
struct BState {
  int I = 0, M = 0, max_ir = 0;
  int32_t *valid, *act, *on, *run, *stop, *ir_it, *kq, *gm_it, *m_of;
  double *bnorm, *relres, *r_prev, *scale, *rho, *dn, *xn, *rn, *fac;
  unsigned long long* t;  // device wall clock when each instance stopped (0: not stopped), then the start
  double *H, *g, *cs, *sn, *y;  // H: (M + 1) x M per instance, column-major; g: M + 1; cs, sn: M; y: M x I
  int32_t* flags;
  // refinement iteration i (1 .. max_ir; max_ir + 1 unused), cycle c (0:
  // the first PrecondGmres, i: iteration i), step k of cycle c
  __host__ __device__ int32_t* act_flag(int i) const { return flags + i; }
  __host__ __device__ int32_t* cyc_flag(int c) const { return flags + max_ir + 2 + c * (M + 1); }
  __host__ __device__ int32_t* step_flag(int c, int k) const { return cyc_flag(c) + 1 + k; }
  __host__ __device__ int nflags() const { return max_ir + 2 + (max_ir + 1) * (M + 1); }
};

__device__ __forceinline__ void raise_flag(int32_t* f) { __hip_atomic_store(f, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// fills in place of hipMemsetAsync: captured graphs replay kernels reliably
__global__ void k_bfill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t < n) p[t] = v;
}
__global__ void k_bfill_f64(double* __restrict__ p, int64_t n, double v, const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t < n) p[t] = v;
}
__global__ void k_bcopy(const double* __restrict__ src, double* __restrict__ dst, int64_t n,
                        const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// start of a solve: every instance's refinement state, and the start time
__global__ void k_bir_start(BState s) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0) s.t[s.I] = wall_clock64();
  if (q >= s.I) return;
  s.act[q] = 0;
  s.stop[q] = 0;
  s.ir_it[q] = 0;
  s.gm_it[q] = 0;
  s.relres[q] = 0.0;
  s.r_prev[q] = DBL_MAX;
  s.t[q] = 0;
}

// after the first PrecondGmres and ||b||: an instance with a non-finite
// right-hand side stops with reason 4, the others refine
__global__ void k_bir_init(BState s) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I || !s.valid[q]) return;
  if (!isfinite(s.bnorm[q])) {
    s.stop[q] = 4;
    s.relres[q] = __builtin_nan("");
    return;
  }
  s.act[q] = 1;
  raise_flag(s.act_flag(1));
}

// refinement iteration i, before its cycle: an exactly zero residual means
// x is exact (stop reason 3); fac = 1 / ||r||
__global__ void k_bir_scale(BState s, int i) {
  if (*s.act_flag(i) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I) return;
  if (!s.act[q]) {
    s.fac[q] = 0.0;
    return;
  }
  if (s.scale[q] == 0.0) {
    s.stop[q] = 3;
    s.ir_it[q] = i;
    s.relres[q] = 0.0;
    s.act[q] = 0;
    s.t[q] = wall_clock64();
    s.fac[q] = 0.0;
    return;
  }
  s.fac[q] = 1.0 / s.scale[q];
}

__global__ void k_bir_unscale(BState s, int i) {
  if (*s.act_flag(i) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < s.I) s.fac[q] = s.act[q] ? s.scale[q] : 0.0;
}

// the stopping rules after the update of iteration i (gmres_ir.hpp:575-614)
__global__ void k_bir_check(BState s, int i, double tol) {
  if (*s.act_flag(i) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I || !s.act[q]) return;
  const double r_norm = s.rn[q] / s.bnorm[q];
  s.ir_it[q] = i;
  s.relres[q] = r_norm;
  int stop = 0;
  if (!isfinite(r_norm) || !isfinite(s.dn[q])) stop = 4;
  else if (s.dn[q] <= s.xn[q] * (10 * DBL_EPSILON)) stop = 1;
  else if (fabs(r_norm - s.r_prev[q]) < tol) stop = 2;
  if (stop == 0) {
    s.r_prev[q] = r_norm;
    raise_flag(s.act_flag(i + 1));
    return;
  }
  s.stop[q] = stop;
  s.act[q] = 0;
  s.t[q] = wall_clock64();
}

// no refinement iterations: the residual of the first solve
__global__ void k_bir_relres(BState s) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < s.I) s.relres[q] = s.rn[q] / s.bnorm[q];
}

// cycle c: the instances of the mask whose right-hand side is not below the
// tolerance iterate (gmres_ir.hpp:647-650)
__global__ void k_bg_begin(BState s, int c, const int32_t* __restrict__ mask, double tol, const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I) return;
  const int on = mask[q] && !(s.rho[q] < tol);
  s.on[q] = on;
  s.run[q] = on;
  s.kq[q] = 0;
  if (on) raise_flag(s.cyc_flag(c));
}

// after V_0 = M^{-1} b and rho = ||V_0||: V_0 / rho, g = (rho, 0, ..)
__global__ void k_bg_start(BState s, int c) {
  if (*s.cyc_flag(c) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I) return;
  s.fac[q] = s.on[q] ? 1.0 / s.rho[q] : 0.0;
  double* g = s.g + static_cast<int64_t>(q) * (s.M + 1);
  g[0] = s.rho[q];
  for (int j = 1; j <= s.M; ++j) g[j] = 0.0;
}

// step k of cycle c: the instances still iterating and within their cycle
// length (min(max_gmres, n_s), gmres_ir.hpp:643)
__global__ void k_bg_step(BState s, int c, int k) {
  if (*s.cyc_flag(c) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I) return;
  const int run = s.run[q] && k < s.m_of[q];
  s.run[q] = run;
  if (run) raise_flag(s.step_flag(c, k));
}

// Hessenberg column k of every iterating instance from the two CGS passes
// and the norm, the earlier rotations applied, a new rotation and g
// (gmres_ir.hpp:699-730); fac = 1 / H(k + 1, k) for the new basis vector
__global__ void k_bg_givens(BState s, int c, int k, const double* __restrict__ hc, const double* __restrict__ h2,
                            double tol) {
  if (*s.step_flag(c, k) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I) return;
  if (!s.run[q]) {
    s.fac[q] = 0.0;
    return;
  }
  const int I = s.I, M = s.M, k1 = k + 1;
  double* H = s.H + static_cast<int64_t>(q) * (M + 1) * M + static_cast<int64_t>(k) * (M + 1);  // column k
  double* g = s.g + static_cast<int64_t>(q) * (M + 1);
  double* cq = s.cs + static_cast<int64_t>(q) * M;
  double* sq = s.sn + static_cast<int64_t>(q) * M;
  for (int j = 0; j <= k; ++j) H[j] = hc[static_cast<int64_t>(j) * I + q] + h2[static_cast<int64_t>(j) * I + q];
  H[k + 1] = hc[static_cast<int64_t>(k1) * I + q];
  s.fac[q] = H[k + 1] != 0 ? 1.0 / H[k + 1] : 1.0;
  for (int j = 0; j < k; ++j) {
    const double w1 = cq[j] * H[j] - sq[j] * H[j + 1];
    const double w2 = sq[j] * H[j] + cq[j] * H[j + 1];
    H[j] = w1;
    H[j + 1] = w2;
  }
  const double nu = dnrm2_pair(H[k], H[k + 1]);
  if (nu != 0) {
    cq[k] = H[k] / nu;
    sq[k] = -H[k + 1] / nu;
    H[k] = cq[k] * H[k] - sq[k] * H[k + 1];
    H[k + 1] = 0.0;
    const double w1 = cq[k] * g[k] - sq[k] * g[k + 1];
    const double w2 = sq[k] * g[k] + cq[k] * g[k + 1];
    g[k] = w1;
    g[k + 1] = w2;
  }
  s.kq[q] = k + 1;
  s.run[q] = fabs(g[k + 1]) > tol;
}

// back-substitution of every instance (gmres_ir.hpp:731-739): y[j I + q]
__global__ void k_bg_backsolve(BState s, int c) {
  if (*s.cyc_flag(c) == 0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.I) return;
  const int I = s.I, M = s.M, k = s.kq[q];
  s.gm_it[q] += k;
  const double* H = s.H + static_cast<int64_t>(q) * (M + 1) * M;
  const double* g = s.g + static_cast<int64_t>(q) * (M + 1);
  for (int i = k; i-- > 0;) {
    double yi = g[i];
    for (int j = i + 1; j < k; ++j) yi -= H[static_cast<int64_t>(j) * (M + 1) + i] * s.y[static_cast<int64_t>(j) * I + q];
    s.y[static_cast<int64_t>(i) * I + q] = yi / H[static_cast<int64_t>(i) * (M + 1) + i];
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

void GmresLdl::flush_staged() {
  if (staged_.empty()) return;
  // 256-byte aligned places in the pinned buffer
  std::vector<size_t> off(staged_.size() + 1, 0);
  for (size_t i = 0; i < staged_.size(); ++i) off[i + 1] = off[i] + (staged_[i].bytes + 255) / 256 * 256;
  const size_t total = off.back();
  if (total > stage_cap_) {
    MPIR_HIP(hipStreamSynchronize(stream_));  // no copy may still read the old buffer
    if (stage_pinned_) MPIR_HIP(hipHostFree(stage_pinned_));
    stage_pinned_ = nullptr;
    stage_cap_ = 0;
    const size_t cap = std::max<size_t>(total + total / 4, size_t(1) << 24);
    MPIR_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage_pinned_), cap, hipHostMallocDefault));
    stage_cap_ = cap;
  }
  // host copies in 4 MB pieces on up to 8 threads
  constexpr size_t kPiece = size_t(4) << 20;
  std::vector<std::pair<size_t, size_t>> pieces;  // (array, byte offset)
  for (size_t i = 0; i < staged_.size(); ++i)
    for (size_t b = 0; b < staged_[i].bytes; b += kPiece) pieces.push_back({i, b});
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t q; (q = next.fetch_add(1)) < pieces.size();) {
      const Staged& e = staged_[pieces[q].first];
      const size_t b = pieces[q].second, len = std::min(kPiece, e.bytes - b);
      std::memcpy(stage_pinned_ + off[pieces[q].first] + b, static_cast<const char*>(e.src) + b, len);
    }
  };
  const int nt = static_cast<int>(std::min<size_t>(8, std::max<size_t>(1, pieces.size())));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  for (size_t i = 0; i < staged_.size(); ++i)
    MPIR_HIP(hipMemcpyAsync(staged_[i].dst, stage_pinned_ + off[i], staged_[i].bytes, hipMemcpyHostToDevice, stream_));
  staged_.clear();
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
  const double t_join = now_ms();
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
  // every upload below goes through the pinned staging buffer
  staging_ = true;
  struct StageOff {
    GmresLdl* g;
    ~StageOff() {
      g->staging_ = false;
      g->staged_.clear();
    }
  } stage_off{this};
  upload_plan();
  up(d_csr_p_, B->csr_p);
  up(d_csr_j_, B->csr_j);
  up(d_csr_src_, B->csr_src);
  d_csr_x_.alloc(std::max<size_t>(1, B->csr_j.size()) * sizeof(double));
  up(d_entry_col_, B->ecol);
  up(d_entry_row_, B->erow);
  up(d_diag_pos_, B->diag_pos);
  up(d_perm_, perm_);
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
  up(d_b_sys_of_row_, B->sys_of_row);
  up(d_b_col_base_, B->col_base);
  up(d_b_lperm_, B->local_perm);
  up(d_b_ch_start_, ch_start);
  up(d_b_ch_len_, ch_len);
  up(d_b_sys_ch_, sys_ch);
  d_b_piv_.alloc(2 * sizeof(int32_t) * B->nsys);
  alloc_numeric();
  flush_staged();
  staging_ = false;
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
  bstats_.join_ms = t_join - t0;
  bstats_.upload_ms = stats_.analyze_ms - bstats_.join_ms;
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

// Layout of the per-instance state in d_bs_ (I instances, cycles of M
// steps, max_ir refinement iterations), in 256-byte aligned areas from
// `base`; *bytes (if set) receives the total.
BState bstate_layout(void* base, int I, int M, int max_ir, size_t* bytes = nullptr) {
  BState s;
  s.I = I;
  s.M = M;
  s.max_ir = max_ir;
  size_t off = 0;
  auto take = [&](size_t n) {
    const size_t at = off;
    off += (n + 255) / 256 * 256;
    return static_cast<char*>(base) + at;
  };
  const size_t ni = static_cast<size_t>(I) * sizeof(int32_t), nd = static_cast<size_t>(I) * sizeof(double);
  const size_t m1 = static_cast<size_t>(std::max(M, 1));
  int32_t** ints[] = {&s.valid, &s.act, &s.on, &s.run, &s.stop, &s.ir_it, &s.kq, &s.gm_it, &s.m_of};
  for (int32_t** v : ints) *v = reinterpret_cast<int32_t*>(take(ni));
  double** dbl[] = {&s.bnorm, &s.relres, &s.r_prev, &s.scale, &s.rho, &s.dn, &s.xn, &s.rn, &s.fac};
  for (double** v : dbl) *v = reinterpret_cast<double*>(take(nd));
  s.t = reinterpret_cast<unsigned long long*>(take((static_cast<size_t>(I) + 1) * sizeof(unsigned long long)));
  s.H = reinterpret_cast<double*>(take(nd * (M + 1) * m1));
  s.g = reinterpret_cast<double*>(take(nd * (M + 1)));
  s.cs = reinterpret_cast<double*>(take(nd * m1));
  s.sn = reinterpret_cast<double*>(take(nd * m1));
  s.y = reinterpret_cast<double*>(take(nd * m1));
  s.flags = reinterpret_cast<int32_t*>(take(static_cast<size_t>(s.nflags()) * sizeof(int32_t)));
  if (bytes) *bytes = off;
  return s;
}

// GmresLDLIR::Solve (gmres_ir.hpp:534-634) for every (system, right-hand
// side) instance in lock step, with PrecondGmres (gmres_ir.hpp:636-741) as
// the cycles: cycle 0 solves with b, cycle i (refinement iteration i) with
// the scaled residual. Only kernels are launched (fills in place of
// memsets), so the sequence can be captured as it stands.
void GmresLdl::batch_enqueue(const BatchRun& br) {
  hipStream_t st = stream_;
  const int R = br.R, M = br.M, S = batch_->nsys, I = R * S;
  const int64_t N = n_, RN = static_cast<int64_t>(R) * N, ldv = RN;
  const int maxir = static_cast<int>(max_ir_);
  Seg g{n_, R, S, b_nch_, d_b_sys_of_row_.as<int32_t>(), d_b_ch_start_.as<int32_t>(), d_b_ch_len_.as<int32_t>(),
        d_b_sys_ch_.as<int32_t>()};
  const BState bs = bstate_layout(d_bs_.get(), I, M, maxir);
  double* V = d_V_.as<double>();
  double* hc = d_hcol_.as<double>();
  double* h2 = d_small_.as<double>();
  double* part = d_red_.as<double>();
  const double tol = tol_;
  const dim3 gi(nblk(I)), gv(nblk(RN)), gch(b_nch_, R);
  // per-instance reductions (op 1: 2-norm, 2: inf-norm) into out[q]
  auto bred = [&](const double* v, int op, double* out, const int* gate) {
    hipLaunchKernelGGL(k_bred_partial, gch, dim3(kBT), 0, st, g, v, nullptr, op, part, gate);
    hipLaunchKernelGGL(k_bred_final, gi, dim3(256), 0, st, g, part, op, out, gate);
  };
  auto scale = [&](const double* x, double* out, const int* gate) {
    hipLaunchKernelGGL(k_bscale, gv, dim3(256), 0, st, g, x, bs.fac, out, gate);
  };
  auto precond = [&](double* x, const int* gate) {
    double* u = d_uvec_.as<double>();
    if (uf_ == kF64)
      ldl_solve_multi<double, double, double>(solve_args<double, double>(d_panel_.get()), dsp_, x, R, u, st, gate);
    else if (uf_ == kF32)
      ldl_solve_multi<float, float, double>(solve_args<float, float>(d_panel_.get()), dsp_, x, R, u, st, gate);
    else if (uf_ == kF16)
      ldl_solve_multi<__half, float, double>(solve_args<__half, float>(d_panel16_.get()), dsp_, x, R, u, st, gate);
    else
      ldl_solve_multi<__hip_bfloat16, float, double>(solve_args<__hip_bfloat16, float>(d_panel16_.get()), dsp_, x, R,
                                                     u, st, gate);
  };
  auto product = [&](const double* x, const double* b, double* y, const int* gate) {
    CsrDev A{n_, d_csr_p_.as<int32_t>(), d_csr_j_.as<int32_t>(), d_csr_x_.as<double>(), gate};
    for (int r = 0; r < R; ++r) {
      const int64_t o = static_cast<int64_t>(r) * N;
      if (b) residual_f64(A, x + o, b + o, y + o, st);
      else spmv_f64(A, x + o, y + o, st);
    }
  };
  // one PrecondGmres cycle for the instances of mask: res = V y
  auto cycle = [&](int c, const int32_t* mask, const int* gate, const double* b, double* res) {
    hipLaunchKernelGGL(k_bfill_f64, gv, dim3(256), 0, st, res, RN, 0.0, gate);
    bred(b, 1, bs.rho, gate);
    hipLaunchKernelGGL(k_bg_begin, gi, dim3(256), 0, st, bs, c, mask, tol, gate);
    const int* cg = bs.cyc_flag(c);
    hipLaunchKernelGGL(k_bcopy, gv, dim3(256), 0, st, b, V, RN, cg);
    precond(V, cg);
    bred(V, 1, bs.rho, cg);
    hipLaunchKernelGGL(k_bg_start, gi, dim3(256), 0, st, bs, c);
    scale(V, V, cg);
    for (int k = 0; k < M; ++k) {
      hipLaunchKernelGGL(k_bg_step, gi, dim3(256), 0, st, bs, c, k);
      const int* sg = bs.step_flag(c, k);
      double* w = V + ldv * (k + 1);
      product(V + ldv * k, nullptr, w, sg);
      precond(w, sg);
      const int k1 = k + 1;
      for (int pass = 0; pass < 2; ++pass) {
        double* hp = pass == 0 ? hc : h2;
        hipLaunchKernelGGL(k_bmdot_partial, gch, dim3(kBT), 0, st, g, V, ldv, k1, w, part, sg);
        hipLaunchKernelGGL(k_bmdot_final, dim3(nblk(static_cast<int64_t>(k1) * I)), dim3(256), 0, st, g, part, k1, hp, 0,
                           sg);
        hipLaunchKernelGGL(k_bsub_basis, gv, dim3(256), 0, st, g, V, ldv, k1, hp, w, sg);
      }
      bred(w, 1, hc + static_cast<int64_t>(k1) * I, sg);
      hipLaunchKernelGGL(k_bg_givens, gi, dim3(256), 0, st, bs, c, k, hc, h2, tol);
      scale(w, w, sg);
    }
    hipLaunchKernelGGL(k_bg_backsolve, gi, dim3(256), 0, st, bs, c);
    hipLaunchKernelGGL(k_bcombine, gv, dim3(256), 0, st, g, V, ldv, bs.kq, bs.y, res, cg);
  };
  hipLaunchKernelGGL(k_bfill_i32, dim3(nblk(bs.nflags())), dim3(256), 0, st, bs.flags, static_cast<int64_t>(bs.nflags()),
                     0);
  hipLaunchKernelGGL(k_bir_start, gi, dim3(256), 0, st, bs);
  hipLaunchKernelGGL(k_bperm_in, gv, dim3(256), 0, st, g, br.io_in, d_b_off_.as<int64_t>(), d_b_nrhs_.as<int32_t>(),
                     d_b_col_base_.as<int32_t>(), d_b_lperm_.as<int32_t>(), d_ad_.as<double>(), br.vb);
  cycle(0, bs.valid, nullptr, br.vb, br.vx);
  product(br.vx, br.vb, br.vr, nullptr);
  bred(br.vb, 2, bs.bnorm, nullptr);
  hipLaunchKernelGGL(k_bir_init, gi, dim3(256), 0, st, bs);
  for (int i = 1; i <= maxir; ++i) {
    const int* ag = bs.act_flag(i);
    bred(br.vr, 2, bs.scale, ag);
    hipLaunchKernelGGL(k_bir_scale, gi, dim3(256), 0, st, bs, i);
    scale(br.vr, br.vt, ag);
    cycle(i, bs.act, ag, br.vt, br.vd);
    hipLaunchKernelGGL(k_bir_unscale, gi, dim3(256), 0, st, bs, i);
    scale(br.vd, br.vd, ag);
    bred(br.vd, 2, bs.dn, ag);
    bred(br.vx, 2, bs.xn, ag);
    hipLaunchKernelGGL(k_badd, gv, dim3(256), 0, st, g, br.vx, br.vd, bs.act, ag);
    product(br.vx, br.vb, br.vr, ag);
    bred(br.vr, 2, bs.rn, ag);
    hipLaunchKernelGGL(k_bir_check, gi, dim3(256), 0, st, bs, i, tol);
  }
  if (maxir == 0) {
    bred(br.vr, 2, bs.rn, nullptr);
    hipLaunchKernelGGL(k_bir_relres, gi, dim3(256), 0, st, bs);
  }
  hipLaunchKernelGGL(k_bperm_out, gv, dim3(256), 0, st, g, br.vx, d_b_off_.as<int64_t>(), d_b_nrhs_.as<int32_t>(),
                     d_b_col_base_.as<int32_t>(), d_b_lperm_.as<int32_t>(), d_ad_.as<double>(), br.io_out);
  MPIR_HIP(hipGetLastError());
}

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
  // instances q = r S + s; the cycle length of each is min(max_gmres, n_s)
  // (gmres_ir.hpp:643), M the longest of the valid ones
  std::vector<int32_t> valid_m(2 * static_cast<size_t>(I), 0);
  int M = 0;
  for (int q = 0; q < I; ++q) {
    const int s = q % S;
    valid_m[q] = (q / S) < nrhs[s];
    valid_m[I + q] = static_cast<int32_t>(std::min<int64_t>(max_gmres_, batch_n(s)));
    if (valid_m[q]) M = std::max(M, valid_m[I + q]);
  }
  MPIR_CHECK(M + 2 <= 4096, "GMRES iteration limit above 4094");
  // per-system right-hand sides packed into one device array (ld = n_s)
  std::vector<int64_t> off(S + 1, 0);
  for (int s = 0; s < S; ++s) off[s + 1] = off[s] + static_cast<int64_t>(nrhs[s]) * batch_n(s);
  d_b_vec_.ensure(static_cast<size_t>(5 * RN + 2 * off[S]) * sizeof(double));
  BatchRun br;
  br.R = R;
  br.M = M;
  br.vb = d_b_vec_.as<double>();
  br.vx = br.vb + RN;
  br.vr = br.vx + RN;
  br.vd = br.vr + RN;
  br.vt = br.vd + RN;
  br.io_in = br.vt + RN;
  br.io_out = br.io_in + off[S];
  // every buffer of the sequence exists before it is launched or captured
  d_V_.ensure(static_cast<size_t>(M + 1) * RN * sizeof(double));
  d_hcol_.ensure(static_cast<size_t>(M + 2) * I * sizeof(double));
  d_small_.ensure(static_cast<size_t>(M + 2) * I * sizeof(double));
  d_red_.ensure(static_cast<size_t>(std::max(M, 2)) * R * b_nch_ * sizeof(double));
  size_t state_bytes = 0;
  (void)bstate_layout(nullptr, I, M, static_cast<int>(max_ir_), &state_bytes);
  d_bs_.ensure(state_bytes);
  set_solve_slots(std::min(R, 8));
  const BState bs = bstate_layout(d_bs_.get(), I, M, static_cast<int>(max_ir_));
  for (int s = 0; s < S; ++s)
    if (nrhs[s] > 0)
      MPIR_HIP(hipMemcpyAsync(br.io_in + off[s], b[s], (off[s + 1] - off[s]) * sizeof(double),
                              device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  std::vector<int32_t> nr(nrhs, nrhs + S);
  upload_vec(d_b_off_, off, st);
  upload_vec(d_b_nrhs_, nr, st);
  MPIR_HIP(hipMemcpyAsync(bs.valid, valid_m.data(), I * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MPIR_HIP(hipMemcpyAsync(bs.m_of, valid_m.data() + I, I * sizeof(int32_t), hipMemcpyHostToDevice, st));
  // the key of this solve's launch sequence: every scalar and address its
  // kernels take
  std::vector<uint64_t> key = {static_cast<uint64_t>(R), static_cast<uint64_t>(S), static_cast<uint64_t>(M),
                               static_cast<uint64_t>(max_ir_), static_cast<uint64_t>(schedule_),
                               static_cast<uint64_t>(uf_), plan_epoch_, static_cast<uint64_t>(n_)};
  uint64_t tol_bits = 0;
  std::memcpy(&tol_bits, &tol_, sizeof tol_bits);
  key.push_back(tol_bits);
  for (const DevBuf* d : {&d_b_vec_, &d_V_, &d_hcol_, &d_small_, &d_red_, &d_bs_, &d_uvec_, &d_scratch_, &d_panel_,
                          &d_panel16_, &d_D_, &d_Dinv_, &d_counters_, &d_serr_, &d_csr_p_, &d_csr_j_, &d_csr_x_,
                          &d_ad_, &d_b_off_, &d_b_nrhs_, &d_b_col_base_, &d_b_lperm_, &d_b_sys_of_row_,
                          &d_b_ch_start_, &d_b_ch_len_, &d_b_sys_ch_})
    key.push_back(reinterpret_cast<uint64_t>(d->get()));
  key.push_back(static_cast<uint64_t>(off[S]));
  bool replay = false;
  const bool cached = bgraph_ != nullptr && key == bgraph_key_;
  if (batch_graph_ && (cached || key == blast_key_)) {
    if (!cached) {
      if (bgraph_) {
        (void)hipGraphExecDestroy(bgraph_);
        bgraph_ = nullptr;
      }
      hipGraph_t graph = nullptr;
      MPIR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        batch_enqueue(br);
      } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      MPIR_HIP(hipStreamEndCapture(st, &graph));
      size_t nodes = 0;
      MPIR_HIP(hipGraphGetNodes(graph, nullptr, &nodes));
      const hipError_t e = hipGraphInstantiate(&bgraph_, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      MPIR_HIP(e);
      bgraph_nodes_ = static_cast<int32_t>(nodes);
      bgraph_key_ = key;
    }
    MPIR_HIP(hipGraphLaunch(bgraph_, st));
    replay = true;
  } else {
    batch_enqueue(br);
  }
  blast_key_ = key;
  for (int s = 0; s < S; ++s)
    if (nrhs[s] > 0)
      MPIR_HIP(hipMemcpyAsync(x[s], br.io_out + off[s], (off[s + 1] - off[s]) * sizeof(double),
                              device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
  // the per-instance outcome
  std::vector<int32_t> h_stop(I), h_ir(I), h_gm(I), h_act(I), h_flags(bs.nflags());
  std::vector<double> h_rel(I);
  std::vector<unsigned long long> h_t(static_cast<size_t>(I) + 1);
  auto get = [&](void* dst, const void* src, size_t bytes) {
    MPIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  };
  get(h_stop.data(), bs.stop, I * sizeof(int32_t));
  get(h_ir.data(), bs.ir_it, I * sizeof(int32_t));
  get(h_gm.data(), bs.gm_it, I * sizeof(int32_t));
  get(h_act.data(), bs.act, I * sizeof(int32_t));
  get(h_flags.data(), bs.flags, h_flags.size() * sizeof(int32_t));
  get(h_rel.data(), bs.relres, I * sizeof(double));
  get(h_t.data(), bs.t, h_t.size() * sizeof(unsigned long long));
  sync_check();
  const double t_end = now_ms() - t_start;
  int dev = 0, khz = 0;
  MPIR_HIP(hipGetDevice(&dev));
  MPIR_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  // lock-step counts from the flags
  int32_t steps = 0, cycles = 0;
  const int maxir = static_cast<int>(max_ir_);
  for (int c = 0; c <= maxir; ++c) {
    cycles += h_flags[maxir + 2 + c * (M + 1)] ? 1 : 0;
    for (int k = 0; k < M; ++k) steps += h_flags[maxir + 2 + c * (M + 1) + 1 + k] ? 1 : 0;
  }
  bstats_.max_rhs = R;
  bstats_.arnoldi_steps = steps;
  bstats_.precond_sweeps = cycles + steps;
  bstats_.graph_nodes = replay ? bgraph_nodes_ : 0;
  stats_.precond_applications += static_cast<int64_t>(R) * (cycles + steps);
  // per-instance statistics in system-major order (s * R + r)
  bstats_.ir_iterations.assign(I, 0);
  bstats_.gmres_iterations.assign(I, 0);
  bstats_.stop_reason.assign(I, 0);
  bstats_.final_rel_residual.assign(I, 0.0);
  bstats_.done_ms.assign(I, 0.0);
  for (int q = 0; q < I; ++q) {
    const int r = q / S, s = q % S, o = s * R + r;
    bstats_.ir_iterations[o] = h_ir[q];
    bstats_.gmres_iterations[o] = h_gm[q];
    bstats_.stop_reason[o] = h_stop[q];
    bstats_.final_rel_residual[o] = h_rel[q];
    if (h_t[q] != 0 && khz > 0) bstats_.done_ms[o] = static_cast<double>(h_t[q] - h_t[I]) / khz;
    else if (valid_m[q] && h_act[q]) bstats_.done_ms[o] = t_end;  // iteration limit
  }
  bstats_.solve_ms = now_ms() - t_start;
  stats_.solve_ms = bstats_.solve_ms;
}

}  // namespace mpir
