// GMRES-IR on the device: vector kernels and the driver.
//
// Algorithm (yex33/MPIR include/MPIR/gmres_ir.hpp, LDL^T variant of the MIS,
// docs/Design/SoftDetailedDes/MIS.tex:505-568):
//   Compute: AD_i = 1/sqrt|a_ii| when |a_ii| > eps(UF) (gmres_ir.hpp:347-363),
//            K_s = AD K AD, fill-reducing order, LDL^T of K_s in UF.
//   Solve  : b_s = b o AD; x = PrecondGmres(b_s); r = b_s - K_s x (FP64);
//            repeat: scale = ||r||_inf; d = PrecondGmres(r * (1/scale)) * scale;
//            x += d; r = b_s - K_s x; stop when ||d|| <= 10 eps ||x_old|| or
//            | ||r||/||b_s|| - previous | < tol (gmres_ir.hpp:534-634);
//            return x o AD.
//   PrecondGmres: left-preconditioned GMRES without restart, M^{-1} =
//            (L D L^T)^{-1}, Givens rotations c = h_kk/nu, s = -h_k+1,k/nu,
//            back substitution, x = V y (gmres_ir.hpp:636-741). The basis is
//            orthogonalised by classical Gram-Schmidt applied twice (CGS2)
//            instead of modified Gram-Schmidt with a conditional second pass.
// Vectors stay on the device in the fill-reducing order; only the Hessenberg
// column and the norms that drive the control flow come back to the host.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <thread>

#include "host/sparse.hpp"
#include "kernels/norms.hpp"
#include "solver.hpp"

namespace mpir {

enum GmDoubleSlot : int { kGsRhoB = 0, kGsRho, kGsInv, kGsHn, kGsHead = 8 };
enum GmIntSlot : int { kGiCall = 0, kGiActive, kGiK, kGiTotalK, kGiTotalApps, kGiIrGate, kGiCount = 8 };
enum IrSlot : int { kIrBnorm = 0, kIrScale, kIrDnorm, kIrXnorm, kIrRraw, kIrRprev, kIrR0, kIrCount = 8 };
// mapped host words (doubles): the last finished Arnoldi step's ordinal and
// its gate, then the refinement record
enum MirrorSlot : int {
  kMirStep = 4000, kMirActive, kMirIrSeq = 4004,
  kMirIr = 4008,  // iteration, stop reason, r_norm, d_norm, gmres steps, M^-1 applications, ||b_s||_inf, ||r_0||_inf
};

struct GmDev {
  double* gs;  // kGsHead scalars, then g (m + 1), cs (m), sn (m), y (m), H ((m + 1) m)
  int* gi;
  int m;
  __host__ __device__ double* g() const { return gs + kGsHead; }
  __host__ __device__ double* cs() const { return g() + m + 1; }
  __host__ __device__ double* sn() const { return cs() + m; }
  __host__ __device__ double* y() const { return sn() + m; }
  __host__ __device__ double* H() const { return y() + m; }
};

namespace {

constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) r += sh[i];
  __syncthreads();
  return r;
}

// NaN-propagating maximum over the block (kernels/norms.hpp)
__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_down(v, off));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) r = nan_max(r, sh[i]);
  __syncthreads();
  return r;
}

// op 0: sum x_i y_i; op 1: the 2-norm as a scaled sum of squares, one
// (scale, ssq) pair per block in part[2b], part[2b + 1]; op 2: max |x_i|
__global__ void __launch_bounds__(kRedThreads) k_reduce_partial(int n, const double* __restrict__ x,
                                                                const double* __restrict__ y, int op,
                                                                double* __restrict__ part,
                                                                const int* __restrict__ gate = nullptr) {
  __shared__ double sh[2 * kRedThreads / 64];
  if (gate != nullptr && *gate == 0) return;
  if (op == 1) {
    Ssq acc = ssq_init();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc = ssq_add(acc, x[i]);
    const Ssq r = block_ssq(acc, sh);
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = r.scale;
      part[2 * blockIdx.x + 1] = r.ssq;
    }
    return;
  }
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double v = x[i];
    if (op == 0) acc += v * y[i];
    else acc = nan_max(acc, fabs(v));
  }
  const double r = op == 2 ? block_max(acc, sh) : block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// out[slot] = reduction of the partials (op 1: the 2-norm from the pairs)
// With mirror (pinned host memory, device-mapped): mirror[0 .. nhead) =
// head[0 .. nhead) and mirror[nhead] = the result, so the host reads them
// after the stream's work completes without a copy.
__global__ void __launch_bounds__(kRedThreads) k_reduce_final(const double* __restrict__ part, int np, int op,
                                                              double* __restrict__ out, const double* __restrict__ head,
                                                              int nhead, double* __restrict__ mirror,
                                                              const int* __restrict__ gate = nullptr) {
  __shared__ double sh[2 * kRedThreads / 64];
  if (gate != nullptr && *gate == 0) return;
  double v;
  if (op == 1) {
    Ssq acc = ssq_init();
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc = ssq_merge(acc, Ssq{part[2 * i], part[2 * i + 1]});
    v = ssq_value(block_ssq(acc, sh));
  } else {
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc = op == 2 ? nan_max(acc, part[i]) : acc + part[i];
    v = op == 2 ? block_max(acc, sh) : block_sum(acc, sh);
  }
  if (threadIdx.x == 0) *out = v;
  if (mirror != nullptr) {
    for (int i = threadIdx.x; i < nhead; i += blockDim.x) mirror[i] = head[i];
    if (threadIdx.x == 0) mirror[nhead] = v;
    __threadfence_system();
  }
}

// partial dots of w against k1 basis vectors: part[j * gridDim.x + block]
// (gate: the kernels of an Arnoldi step return at once when *gate == 0)
__global__ void __launch_bounds__(kRedThreads) k_mdot_partial(int n, const double* __restrict__ V, int64_t ldv, int k1,
                                                              const double* __restrict__ w, double* __restrict__ part,
                                                              const int* __restrict__ gate) {
  __shared__ double sh[kRedThreads / 64];
  if (gate != nullptr && *gate == 0) return;
  for (int j = 0; j < k1; ++j) {
    const double* vj = V + ldv * j;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += vj[i] * w[i];
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) part[static_cast<int64_t>(j) * gridDim.x + blockIdx.x] = r;
  }
}

// h[j] (+)= sum of the partials of vector j, one block per j
__global__ void __launch_bounds__(kRedThreads) k_mdot_final(const double* __restrict__ part, int np, double* __restrict__ h,
                                                            int accumulate, const int* __restrict__ gate) {
  __shared__ double sh[kRedThreads / 64];
  if (gate != nullptr && *gate == 0) return;
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[static_cast<int64_t>(j) * np + i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) h[j] = accumulate ? h[j] + r : r;
}

// w -= sum_j V_j h_j
__global__ void k_sub_basis(int n, const double* __restrict__ V, int64_t ldv, int k1, const double* __restrict__ h,
                            double* __restrict__ w, const int* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (gate != nullptr && *gate == 0)) return;
  double acc = 0.0;
  for (int j = 0; j < k1; ++j) acc += V[ldv * j + i] * h[j];
  w[i] -= acc;
}

// res = sum_j V_j y_j, j ascending from zero (gmres_ir.hpp:735-739)
__global__ void k_combine(int n, const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ y,
                          double* __restrict__ res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < k; ++j) acc += V[ldv * j + i] * y[j];
  res[i] = acc;
}

// out = x * s where s = scalars[slot] (recip: s = 1 / scalars[slot])
__global__ void k_scale_dev(int n, const double* __restrict__ x, const double* __restrict__ sp, int recip,
                            double* __restrict__ out, const int* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (gate != nullptr && *gate == 0)) return;
  const double s = recip ? 1.0 / *sp : *sp;
  out[i] = x[i] * s;
}

__global__ void k_scale_host(int n, const double* __restrict__ x, double s, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = x[i] * s;
}

__global__ void k_add(int n, double* __restrict__ x, const double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + d[i];
}
__global__ void k_add_gated(int n, double* __restrict__ x, const double* __restrict__ d, const int* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && *gate != 0) x[i] = x[i] + d[i];
}
__global__ void k_copy_gated(int n, const double* __restrict__ src, double* __restrict__ dst,
                             const int* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && *gate != 0) dst[i] = src[i];
}

// ------------------------------------------------------- GMRES-IR on the device
// PrecondGmres (gmres_ir.hpp:636-741) keeps its scalars in device memory: the
// Hessenberg matrix (column-major, ld m + 1), the Givens rotations, g, rho and
// the step count, plus gates: a kernel of an Arnoldi step returns at once when
// the iteration has already stopped, so the host queues step k + 1 before
// step k has finished and never waits for a step's scalars. Solve's stopping
// rules (gmres_ir.hpp:599-606) are evaluated by k_ir_check, which hands one
// record per refinement iteration to the host through mapped memory.
__device__ __forceinline__ void mirror_put(double* mirror, int slot, double v) {
  __hip_atomic_store(mirror + slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}


// start of a call: it runs when the caller's gate is open and ||b||_2 is not
// below tol (gmres_ir.hpp:647-650); the first M^-1 application follows
__global__ void k_gm_begin(GmDev d, const int* __restrict__ gate, double tol) {
  const int on = (gate == nullptr || *gate != 0) && !(d.gs[kGsRhoB] < tol);
  d.gi[kGiCall] = on;
  d.gi[kGiActive] = on;
  d.gi[kGiK] = 0;
  if (on) d.gi[kGiTotalApps] += 1;
}

// rho = ||M^-1 b||_2: g = (rho, 0, ...), V_0 scaled by 1 / rho, and the loop
// test k < m && rho > tol for k = 0
__global__ void k_gm_start(GmDev d, double tol) {
  if (!d.gi[kGiActive]) return;
  const double rho = d.gs[kGsRho];
  d.gs[kGsInv] = 1 / rho;
  double* g = d.g();
  g[0] = rho;
  for (int i = 1; i <= d.m; ++i) g[i] = 0.0;
  d.gi[kGiActive] = (d.m > 0 && rho > tol) ? 1 : 0;
}

// Arnoldi step k after its CGS2 passes (column h1 + h2, ||w|| in kGsHn):
// Hessenberg column, the earlier rotations, the new rotation, g, rho and
// the gate of step k + 1 (gmres_ir.hpp:694-723). Publishes the step's
// ordinal and gate to the host whether or not the step ran.
__global__ void k_gm_givens(GmDev d, int k, double tol, const double* __restrict__ h1, const double* __restrict__ h2,
                            double* mirror, double ordinal) {
  if (d.gi[kGiActive]) {
    const int m = d.m;
    double* H = d.H() + static_cast<int64_t>(k) * (m + 1);
    double* g = d.g();
    double* c = d.cs();
    double* s = d.sn();
    for (int j = 0; j <= k; ++j) H[j] = h1[j] + h2[j];
    const double hn = d.gs[kGsHn];
    H[k + 1] = hn;
    d.gs[kGsInv] = hn != 0 ? 1 / hn : 1.0;  // v_{k+1} /= h_{k+1,k} unless it is zero
    for (int j = 0; j < k; ++j) {
      const double w1 = c[j] * H[j] - s[j] * H[j + 1];
      const double w2 = s[j] * H[j] + c[j] * H[j + 1];
      H[j] = w1;
      H[j + 1] = w2;
    }
    const double nu = dnrm2_pair(H[k], H[k + 1]);
    if (nu != 0) {
      c[k] = H[k] / nu;
      s[k] = -H[k + 1] / nu;
      H[k] = c[k] * H[k] - s[k] * H[k + 1];
      H[k + 1] = 0.0;
      const double w1 = c[k] * g[k] - s[k] * g[k + 1];
      const double w2 = s[k] * g[k] + c[k] * g[k + 1];
      g[k] = w1;
      g[k + 1] = w2;
    }
    const double rho = fabs(g[k + 1]);
    d.gs[kGsRho] = rho;
    d.gi[kGiK] = k + 1;
    d.gi[kGiTotalK] += 1;
    d.gi[kGiTotalApps] += 1;
    d.gi[kGiActive] = (k + 1 < m && rho > tol) ? 1 : 0;
  }
  mirror_put(mirror, kMirActive, d.gi[kGiActive]);
  __threadfence_system();
  mirror_put(mirror, kMirStep, ordinal);
  __threadfence_system();
}

// back substitution y = H(0:k, 0:k)^-1 g (gmres_ir.hpp:725-732)
__global__ void k_gm_backsolve(GmDev d) {
  const int k = d.gi[kGiK];
  const int m = d.m;
  const double* H = d.H();
  const double* g = d.g();
  double* y = d.y();
  for (int i = k; i-- > 0;) {
    double v = g[i];
    for (int j = i + 1; j < k; ++j) v -= H[i + static_cast<int64_t>(j) * (m + 1)] * y[j];
    y[i] = v / H[i + static_cast<int64_t>(i) * (m + 1)];
  }
}

// res = sum_{j < k} V_j y_j with the device's step count (zero when k = 0)
__global__ void k_gm_combine(int n, const double* __restrict__ V, int64_t ldv, const int* __restrict__ kp,
                             const double* __restrict__ y, double* __restrict__ res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = *kp;
  double acc = 0.0;
  for (int j = 0; j < k; ++j) acc += V[ldv * j + i] * y[j];
  res[i] = acc;
}

// refinement iteration: the gate opens unless the residual is exactly zero
__global__ void k_ir_gate(const double* __restrict__ ir, int* __restrict__ gi) {
  gi[kGiIrGate] = ir[kIrScale] != 0.0 ? 1 : 0;
}

// Solve's stopping rules after iteration `iter` (gmres_ir.hpp:575-606):
// 1 small update (d_norm <= 10 eps(UW) x_norm, x_norm taken before the
// update), 2 stagnation (|r_norm - r_prev| < tol, absolute), 3 zero residual
// (the gate was closed); the record goes to the host, its sequence number
// `seq` last. iter = 0 records the state after the first PrecondGmres only.
__global__ void k_ir_check(double* __restrict__ ir, const int* __restrict__ gi, int iter, double tol, double eps_uw,
                           double* mirror, double seq) {
  double stop = 0.0, r_norm = 0.0;
  const double bnorm = ir[kIrBnorm];
  if (iter == 0) {
    r_norm = ir[kIrR0] / bnorm;
  } else if (!gi[kGiIrGate]) {
    stop = 3.0;
  } else {
    r_norm = ir[kIrRraw] / bnorm;
    if (ir[kIrDnorm] <= ir[kIrXnorm] * (10 * eps_uw)) stop = 1.0;
    else if (fabs(r_norm - ir[kIrRprev]) < tol) stop = 2.0;
    ir[kIrRprev] = r_norm;
  }
  mirror_put(mirror, kMirIr + 0, iter);
  mirror_put(mirror, kMirIr + 1, stop);
  mirror_put(mirror, kMirIr + 2, r_norm);
  mirror_put(mirror, kMirIr + 3, ir[kIrDnorm]);
  mirror_put(mirror, kMirIr + 4, gi[kGiTotalK]);
  mirror_put(mirror, kMirIr + 5, gi[kGiTotalApps]);
  mirror_put(mirror, kMirIr + 6, bnorm);
  mirror_put(mirror, kMirIr + 7, ir[kIrR0]);
  __threadfence_system();
  mirror_put(mirror, kMirIrSeq, seq);
  __threadfence_system();
}

__global__ void k_ir_reset(double* __restrict__ ir, int* __restrict__ gi) {
  ir[kIrRprev] = DBL_MAX;
  gi[kGiTotalK] = 0;
  gi[kGiTotalApps] = 0;
  gi[kGiIrGate] = 1;
}

__global__ void k_compute_ad(int n, const double* __restrict__ ax, const int32_t* __restrict__ diag_pos, double eps,
                             double* __restrict__ ad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 1.0;
  const int q = diag_pos[i];
  if (q >= 0) {
    const double a = fabs(ax[q]);
    if (a > eps) v = 1.0 / sqrt(a);
  }
  ad[i] = v;
}

__global__ void k_scale_entries(int64_t nnz, const double* __restrict__ ax, const int32_t* __restrict__ row,
                                const int32_t* __restrict__ col, const double* __restrict__ ad, double* __restrict__ axs) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < nnz) axs[q] = ax[q] * (ad[row[q]] * ad[col[q]]);
}

// Symmetric Ruiz equilibration of the full CSR (permuted rows, original
// scale vector ad): rmax[i] = max_j |ad_i k_ij ad_j|; dev[i] = |1 - rmax[i]|.
__global__ void k_ruiz_rowmax(int n, const int32_t* __restrict__ p, const int32_t* __restrict__ j,
                              const double* __restrict__ x, const int32_t* __restrict__ perm,
                              const double* __restrict__ ad, double* __restrict__ rmax, double* __restrict__ dev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ai = ad[perm[i]];
  double mx = 0.0;
  for (int q = p[i]; q < p[i + 1]; ++q) mx = fmax(mx, fabs(ai * x[q] * ad[perm[j[q]]]));
  rmax[i] = mx;
  dev[i] = fabs(1.0 - mx);
}
__global__ void k_ruiz_update(int n, const int32_t* __restrict__ perm, const double* __restrict__ rmax,
                              double* __restrict__ ad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && rmax[i] > 0.0) ad[perm[i]] /= sqrt(rmax[i]);
}
// x[i] = float(x[i]): a vector the reference holds in a float working or
// residual precision, kept in double storage
__global__ void k_round_f32(int64_t n, double* __restrict__ x) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) x[i] = static_cast<double>(static_cast<float>(x[i]));
}
__global__ void k_fill(int n, double v, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}
template <class S, class D>
__global__ void k_convert(int n, const S* __restrict__ src, D* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<D>(src[i]);
}
// x[i] *= s with the product formed in double (tblas.hpp scale)
template <class T>
__global__ void k_scale_by(int n, double s, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = static_cast<T>(static_cast<double>(x[i]) * s);
}
// x[i] = double(c[i]) + x[i]
template <class T>
__global__ void k_add_from(int n, const T* __restrict__ c, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = static_cast<double>(c[i]) + x[i];
}
__global__ void k_gather(int64_t nnz, const double* __restrict__ src, const int32_t* __restrict__ idx,
                         double* __restrict__ out) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < nnz) out[q] = src[idx[q]];
}

// out[k] = b[perm[k]] * ad[perm[k]]  (b o AD, then into the fill-reducing order)
__global__ void k_perm_in(int n, const double* __restrict__ b, const int32_t* __restrict__ perm,
                          const double* __restrict__ ad, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = b[perm[k]] * ad[perm[k]];
}

// x[perm[k]] = xp[k] * ad[perm[k]]
__global__ void k_perm_out(int n, const double* __restrict__ xp, const int32_t* __restrict__ perm,
                           const double* __restrict__ ad, double* __restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[perm[k]] = xp[k] * ad[perm[k]];
}

inline unsigned nblk(int64_t n, int bs = 256) { return static_cast<unsigned>((n + bs - 1) / bs); }

// rounds v to float when prec is f32 (a no-op for f64)
void round_to(int prec, double* v, int64_t n, hipStream_t st) {
  if (prec == kF32 && n > 0) hipLaunchKernelGGL(k_round_f32, dim3(nblk(n)), dim3(256), 0, st, n, v);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
void upload_vec(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (!v.empty()) MPIR_HIP(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// small-vector slots
enum Slot { kSlotNorm = 0, kSlotH = 8 };

}  // namespace

GmresLdl::GmresLdl(int uf, int uw, int ur) : uf_(uf), uw_(uw), ur_(ur) {
  // Refinable<UF,UW,UR> (fp_concepts.hpp:112-113): eps(UF) >= eps(UW) >= eps(UR);
  // the tags are ordered by epsilon (f64 < f32 < f16 < bf16)
  MPIR_CHECK(uf == kF64 || uf == kF32 || uf == kF16 || uf == kBF16, "factor precision must be f64, f32, f16 or bf16");
  MPIR_CHECK((uw == kF64 || uw == kF32) && (ur == kF64 || ur == kF32), "working and residual precision must be f64 or f32");
  MPIR_CHECK(uf >= uw && uw >= ur, "precisions must satisfy eps(UF) >= eps(UW) >= eps(UR) (Refinable)");
  // the solver runs only on a gfx950 device; there is no host fallback
  {
    int count = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1 || hipGetDeviceProperties(&prop, 0) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      throw Error(kErrNoDevice, "no gfx950 device: this library runs the solve path on the GPU only");
  }
  MPIR_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  // coherent: the device writes the norm mirror while the host polls
  MPIR_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_pinned_), 4096 * sizeof(double),
                         hipHostMallocMapped | hipHostMallocCoherent));
  MPIR_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&h_mirror_), h_pinned_, 0));
  // the device publishes records here with increasing sequence numbers from
  // 1: start from zero (the allocation may hold an earlier handle's records)
  std::memset(h_pinned_, 0, 4096 * sizeof(double));
}

// Waits for the stream: polls for up to 200 us (the GMRES loop waits once
// per iteration for a handful of values, and polling answers sooner than a
// blocking wait), then yields the core between polls, so that threads of
// other handles and the host's symbolic work keep running.
void GmresLdl::wait_stream() {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipStreamQuery(stream_);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) MPIR_HIP(e);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) std::this_thread::yield();
  }
}

GmresLdl::~GmresLdl() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (bgraph_) (void)hipGraphExecDestroy(bgraph_);
  if (h_pinned_) (void)hipHostFree(h_pinned_);
  if (b_pinned_) (void)hipHostFree(b_pinned_);
  if (stage_pinned_) (void)hipHostFree(stage_pinned_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void GmresLdl::release_numeric() { factored_ = false; }

void GmresLdl::analyze(int32_t n, const int32_t* Ap, const int32_t* Ai) {
  const double t0 = now_ms();
  MPIR_CHECK(n >= 1, "matrix dimension must be positive");
  MPIR_CHECK(Ap && Ai, "null pattern pointer");
  MPIR_CHECK(Ap[0] == 0, "column pointers must start at 0");
  for (int32_t j = 0; j < n; ++j) {
    MPIR_CHECK(Ap[j + 1] > Ap[j], "empty column in the upper-triangular input");
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) {
      MPIR_CHECK(Ai[q] >= 0 && Ai[q] <= j, "entry below the diagonal or out of range");
      if (q > Ap[j]) MPIR_CHECK(Ai[q] > Ai[q - 1], "row indices must be strictly ascending");
    }
  }
  analyzed_ = factored_ = false;
  batch_.reset();
  n_ = n;
  nnz_ = Ap[n];
  Ap_.assign(Ap, Ap + n + 1);
  Ai_.assign(Ai, Ai + nnz_);
  if (precond_ == kPrecIlu) {
    analyze_ilu(n, Ap, Ai, t0);
    return;
  }
  Analysis an;
  try {
    an = analyze_pattern(n, Ap, Ai, ordering_, relax_);
  } catch (const std::invalid_argument& e) {
    throw Error(kErrInvalid, e.what());
  } catch (const std::overflow_error& e) {
    throw Error(kErrInvalid, e.what());
  }
  perm_ = std::move(an.perm);
  plan_ = std::make_unique<FactorPlan>(std::move(an.plan));
  relaxed_ = an.relaxed;
  const CscUpper& C = an.C;
  const std::vector<int32_t>& val_map = an.val_map;
  Cp_ = C.p;
  Ci_ = C.i;
  parent_ = an.parent;
  Lnz_ = an.Lnz;
  true_nnz_l_ = an.nnz_l;
  true_flops_ = an.flops;
  hipStream_t st = stream_;
  apply_tuning();
  upload_plan();

  // FP64 CSR of the permuted, scaled matrix for the SpMV: entry -> original
  CsrFull F = csr_from_upper(C);
  std::vector<int32_t> src(F.src.size());
  for (size_t q = 0; q < src.size(); ++q) src[q] = val_map[F.src[q]];
  upload_vec(d_csr_p_, F.p, st);
  upload_vec(d_csr_j_, F.j, st);
  upload_vec(d_csr_src_, src, st);
  d_csr_x_.alloc(std::max<size_t>(1, F.j.size()) * sizeof(double));
  // entry rows / columns and diagonal positions of the original pattern
  std::vector<int32_t> ecol(nnz_), diag_pos(n, -1);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = Ap_[j]; q < Ap_[j + 1]; ++q) {
      ecol[q] = j;
      if (Ai_[q] == j) diag_pos[j] = q;
    }
  upload_vec(d_entry_col_, ecol, st);
  upload_vec(d_entry_row_, Ai_, st);
  upload_vec(d_diag_pos_, diag_pos, st);
  upload_vec(d_perm_, perm_, st);

  alloc_numeric();
  MPIR_HIP(hipStreamSynchronize(st));

  const Supernodes& S = plan_->sn;
  stats_ = SolveStats();
  stats_.n = n;
  stats_.nnz = nnz_;
  stats_.nnz_l = true_nnz_l_;
  stats_.nsupernodes = S.ns;
  stats_.nlevels = S.nlevels;
  stats_.max_front = S.max_front;
  stats_.factor_flops = true_flops_;
  stats_.device_bytes = static_cast<int64_t>(d_panel_.bytes() + d_upool_.bytes());
  analyzed_ = true;
  precond_only_ = false;
  stats_.analyze_ms = now_ms() - t0;
}

// Device copies of the dependency-driven factorization's task list,
// counter offsets and counters (empty unless the pool is laid out for it).
void GmresLdl::upload_factor_dag() {
  const FactorPlan& P = *plan_;
  const Supernodes& S = P.sn;
  hipStream_t st = stream_;
  dfp_.use_dag = factor_schedule_ == kScheduleDag;
  dfp_.dag_pool = P.dag_pool;
  dfp_.dag_level = P.fdag_level;
  FactorDagDev& g = dfp_.dag;
  g = FactorDagDev();
  up(d_fdag_, P.fdag);
  up(d_fdag_meta_, P.fdag_meta);
  g.task = d_fdag_.as<FTask>();
  g.ntask = static_cast<int>(P.fdag.size());
  g.meta = d_fdag_meta_.as<int32_t>();
  g.asm_task = dfp_.asm_task;
  g.asm_child_ptr = dfp_.asm_child_ptr;
  g.asm_child = dfp_.asm_child;
  g.ns = S.ns;
  g.nsteps = P.fdag_nsteps;
  const FactorDagIndex X{S.ns, P.fdag_nsteps};
  g.cnt_disp = (X.col_done(P.fdag_ncb) + kDispenserStride - 1) / kDispenserStride * kDispenserStride;
  g.ncounters = g.cnt_disp + kDispenserLines * kDispenserStride;
  d_fcounters_.alloc(static_cast<size_t>(g.ncounters) * sizeof(int));
  g.counters = d_fcounters_.as<int>();
}

void GmresLdl::upload_plan() {
  ++plan_epoch_;
  const FactorPlan& P = *plan_;
  const Supernodes& S = P.sn;
  hipStream_t st = stream_;
  up(d_first_, S.first);
  up(d_nrow_, S.nrow);
  up(d_rows_off_, S.rows_off);
  up(d_rows_, S.rows);
  up(d_relidx_, S.relidx);
  up(d_panel_off_, S.panel_off);
  up(d_upd_off_, P.upd_off);
  up(d_child_ptr_, P.child_ptr);
  up(d_child_list_, P.child_list);
  // the device buffers are kept across re-analyses (a new pattern reuses
  // their memory: freeing would wait for the whole device)
  if (keep_.size() < 40) keep_.resize(40);
  up(keep_[0], P.small_list);
  up(keep_[1], P.big_list);
  up(keep_[2], P.asm_task);
  up(keep_[3], P.pan_task);
  up(keep_[4], P.upd_task);
  up(keep_[5], P.val_dst);
  dfp_ = DeviceFactorPlan();
  dfp_.nlevels = S.nlevels;
  dfp_.panel_entries = P.panel_entries;
  dfp_.small_list = keep_[0].as<int32_t>();
  dfp_.big_list = keep_[1].as<int32_t>();
  dfp_.asm_task = keep_[2].as<Task4>();
  dfp_.pan_task = keep_[3].as<Task4>();
  dfp_.upd_task = keep_[4].as<Task4>();
  dfp_.val_dst = keep_[5].as<int64_t>();
  dfp_.h_small_ptr = P.small_ptr;
  dfp_.h_big_ptr = P.big_ptr;
  dfp_.h_asm_ptr = P.asm_ptr;
  dfp_.h_level_step_ptr = P.level_step_ptr;
  dfp_.h_pan_ptr = P.pan_ptr;
  dfp_.h_upd_ptr = P.upd_ptr;
  dfp_.h_pan_aug_ptr = P.pan_aug_ptr;
  dfp_.n_inv_task = static_cast<int>(P.inv_task.size() / 3);
  dfp_.h_upd_aug_ptr = P.upd_aug_ptr;
  up(keep_[15], P.asm_child_ptr);
  up(keep_[16], P.asm_child);
  dfp_.asm_child_ptr = keep_[15].as<int32_t>();
  dfp_.asm_child = keep_[16].as<int32_t>();
  // split panel steps: a front's Z_kk slot is its place in its level's list
  // of big fronts (levels run one after the other)
  {
    h_zslot_.assign(static_cast<size_t>(std::max(1, S.ns)), -1);
    int32_t max_big = 0;
    for (int32_t l = 0; l < S.nlevels; ++l) {
      max_big = std::max(max_big, P.big_ptr[l + 1] - P.big_ptr[l]);
      for (int32_t q = P.big_ptr[l]; q < P.big_ptr[l + 1]; ++q) h_zslot_[P.big_list[q]] = q - P.big_ptr[l];
    }
    const int64_t slot_bytes = static_cast<int64_t>(kPanelNB) * kPanelNB * (uf_ == kF64 ? 8 : 4);
    dfp_.zbuf_fronts = static_cast<int>(std::min<int64_t>(max_big, kZbufMaxBytes / slot_bytes));
    dfp_.pan_split_min = pan_split_min_;
    up(d_zslot_, h_zslot_);
    // each step's tasks owning a diagonal block (among its real rows: r0 = k0)
    // and the others, as lists of their own for the two launches of a split step
    h_pan_own_.clear();
    h_pan_rest_.clear();
    dfp_.h_pan_own_ptr.assign(1, 0);
    dfp_.h_pan_rest_ptr.assign(1, 0);
    const int32_t nsteps = static_cast<int32_t>(P.pan_ptr.size()) - 1;
    for (int32_t q = 0; q < nsteps; ++q) {
      for (int32_t t = P.pan_ptr[q]; t < P.pan_ptr[q + 1]; ++t) {
        const Task4 tk = P.pan_task[t];
        if (t < P.pan_aug_ptr[q] && tk.b == tk.d * kPanelNB) h_pan_own_.push_back(tk);
        else h_pan_rest_.push_back(tk);
      }
      dfp_.h_pan_own_ptr.push_back(static_cast<int32_t>(h_pan_own_.size()));
      dfp_.h_pan_rest_ptr.push_back(static_cast<int32_t>(h_pan_rest_.size()));
    }
    up(d_pan_own_, h_pan_own_);
    up(d_pan_rest_, h_pan_rest_);
    dfp_.pan_own = d_pan_own_.as<Task4>();
    dfp_.pan_rest = d_pan_rest_.as<Task4>();
    if (dfp_.zbuf_fronts > 0) d_zbuf_.ensure(static_cast<size_t>(dfp_.zbuf_fronts * slot_bytes));
  }
  upload_factor_dag();
  // triangular solves
  dsp_ = DeviceSolvePlan();
  dsp_.schedule = schedule_;
  dsp_.nlevels = S.nlevels;
  dsp_.h_small_ptr = P.small_ptr;
  dsp_.h_fx_ptr = P.fx_ptr;
  dsp_.h_fu_ptr = P.fu_ptr;
  dsp_.h_br_ptr = P.br_ptr;
  dsp_.h_bx_ptr = P.bx_ptr;
  dsp_.h_fm_ptr = P.fm_ptr;
  dsp_.h_bm_ptr = P.bm_ptr;
  SolveDev& g = dsp_.d;
  g.ns = S.ns;
  g.nfwd = static_cast<int>(P.dag_fwd.size());
  g.nbwd = static_cast<int>(P.dag_bwd.size());
  int k = 17;
  auto upk = [&](const auto& v) {
    up(keep_[k], v);
    return keep_[k++].get();
  };
  g.fwd = static_cast<const Task4*>(upk(P.dag_fwd));
  g.bwd = static_cast<const Task4*>(upk(P.dag_bwd));
  g.meta = static_cast<const int32_t*>(upk(P.dag_meta));
  g.small_list = dfp_.small_list;
  g.fx_task = static_cast<const int32_t*>(upk(P.fx_task));
  dfp_.inv_task = static_cast<const int32_t*>(upk(P.inv_task));
  g.fu_task = static_cast<const int32_t*>(upk(P.fu_task));
  g.br_task = static_cast<const int32_t*>(upk(P.br_task));
  g.bx_task = static_cast<const int32_t*>(upk(P.bx_task));
  g.fm_task = static_cast<const int32_t*>(upk(P.fm_task));
  g.bm_task = static_cast<const int32_t*>(upk(P.bm_task));
  g.frow_off = static_cast<const int32_t*>(upk(P.frow_off));
  g.grow_ptr = static_cast<const int32_t*>(upk(P.grow_ptr));
  g.gent = static_cast<const Task4*>(upk(P.gent));
  g.cb_base = static_cast<const int32_t*>(upk(P.cb_base));
  g.ypart_off = static_cast<const int64_t*>(upk(P.ypart_off));
  g.upart_off = static_cast<const int64_t*>(upk(P.upart_off));
  g.rpart_off = static_cast<const int64_t*>(upk(P.rpart_off));
  g.xpart_off = static_cast<const int64_t*>(upk(P.xpart_off));
  MPIR_CHECK(k <= 40, "plan upload slots");
  solve_slots_ = 0;
  set_solve_slots(1);
  // counters: per front, per column block, then the dispensers on their own lines
  g.cnt_bx = kNumSolveCounters * S.ns;
  g.cnt_disp = (g.cnt_bx + P.ncb_total + kDispenserStride - 1) / kDispenserStride * kDispenserStride;
  g.ncounters = g.cnt_disp + 2 * kDispenserLines * kDispenserStride;
  d_counters_.alloc(static_cast<size_t>(g.ncounters) * sizeof(int));
  g.counters = d_counters_.as<int>();
  d_serr_.alloc(sizeof(int));
  MPIR_HIP(hipMemsetAsync(d_serr_.get(), 0, sizeof(int), st));
  g.err = d_serr_.as<int>();
}

// Work areas of the triangular solves for `slots` right-hand sides at once
// (in the widest solve precision): the update vectors and the partial sums
// of every right-hand side, at fixed strides.
void GmresLdl::set_solve_slots(int slots) {
  if (slots <= solve_slots_) return;
  const FactorPlan& P = *plan_;
  SolveDev& g = dsp_.d;
  const int64_t nwork = slots * (P.nypart + P.nupart + P.nrpart + P.nxpart) + 64;
  d_scratch_.alloc(static_cast<size_t>(nwork) * sizeof(double));
  double* wk = d_scratch_.as<double>();
  g.ypart = wk;
  g.upart = wk + slots * P.nypart;
  g.rpart = wk + slots * (P.nypart + P.nupart);
  g.xpart = wk + slots * (P.nypart + P.nupart + P.nrpart);
  g.syp = P.nypart;
  g.sup = P.nupart;
  g.srp = P.nrpart;
  g.sxp = P.nxpart;
  const int64_t rows = std::max<int64_t>(1, static_cast<int64_t>(P.sn.rows.size()));
  d_uvec_.alloc(static_cast<size_t>(slots * rows) * sizeof(double));
  g.su = rows;
  g.sx = n_;
  solve_slots_ = slots;
}

void GmresLdl::alloc_vectors() {
  const int32_t n = n_;
  d_ax_.alloc(std::max<int64_t>(1, nnz_) * sizeof(double));
  d_axs_.alloc(std::max<int64_t>(1, nnz_) * sizeof(double));
  d_ad_.alloc(n * sizeof(double));
  d_err_.alloc(sizeof(int));
  d_cnt_.alloc(sizeof(int));
  d_nonfinite_.alloc(sizeof(int));
  d_b_.alloc(n * sizeof(double));
  d_x_.alloc(n * sizeof(double));
  d_r_.alloc(n * sizeof(double));
  d_d_.alloc(n * sizeof(double));
  d_w_.alloc(n * sizeof(double));
  d_red_.alloc(static_cast<size_t>(kRedBlocks) * 64 * sizeof(double));
  d_small_.alloc(256 * sizeof(double));
}

void GmresLdl::alloc_numeric() {
  const FactorPlan& P = *plan_;
  const int32_t n = n_;
  const size_t tf = uf_ == kF64 ? sizeof(double) : sizeof(float);
  alloc_vectors();
  d_panel_.alloc((std::max<int64_t>(1, P.panel_entries) + kPanelSlack) * tf);
  if (uf_ == kF16 || uf_ == kBF16) d_panel16_.alloc((std::max<int64_t>(1, P.panel_entries) + kPanelSlack) * 2);
  // the slack past the last front holds zeros: the solves' 16-byte tile
  // loads may read into it, and they rely on every value they read being
  // finite
  MPIR_HIP(hipMemsetAsync(d_panel_.as<char>() + P.panel_entries * tf, 0, kPanelSlack * tf, stream_));
  if (uf_ == kF16 || uf_ == kBF16)
    MPIR_HIP(hipMemsetAsync(d_panel16_.as<char>() + P.panel_entries * 2, 0, kPanelSlack * 2, stream_));
  // the assembly's 16-byte loads of a child's update matrix may read up to
  // 3 entries past the last one (those values are not used)
  d_upool_.alloc((std::max<int64_t>(1, P.upd_pool) + kPanelSlack) * tf);
  d_D_.alloc(n * tf);
  d_Dinv_.alloc(n * tf);
}

void GmresLdl::load_factor_csc(int32_t n, const int64_t* Lp, const int32_t* Li, const void* Lx, const void* Dinv) {
  MPIR_CHECK(n >= 1, "matrix dimension must be positive");
  MPIR_CHECK(precond_ == kPrecLdl, "a loaded factor is an L D L^T factor");
  // the plan of the factor pattern is kept while the pattern stays the same
  // (compared entry by entry), so a caller that solves in a loop with the
  // same factor (luir.hpp:165-205) pays only the upload of the values
  const bool same_pattern = precond_only_ && analyzed_ && n == n_ && loaded_Lp_.size() == static_cast<size_t>(n) + 1 &&
                            std::equal(Lp, Lp + n + 1, loaded_Lp_.begin()) &&
                            std::equal(Li, Li + Lp[n], loaded_Li_.begin());
  factored_ = false;
  if (!same_pattern) {
    analyzed_ = false;
    loaded_Lp_.clear();
    loaded_Li_.clear();
    batch_.reset();
    Supernodes S;
    try {
      S = supernodes_from_l(n, Lp, Li);
    } catch (const std::exception& e) {
      throw Error(kErrInvalid, std::string("factor pattern rejected: ") + e.what());
    }
    n_ = n;
    nnz_ = 0;
    perm_.resize(n);
    std::iota(perm_.begin(), perm_.end(), 0);
    plan_ = std::make_unique<FactorPlan>(build_plan_from_supernodes(std::move(S), ftune_, stune_));
    upload_plan();
    alloc_numeric();
    loaded_Lp_.assign(Lp, Lp + n + 1);
    loaded_Li_.assign(Li, Li + Lp[n]);
    ++plan_builds_;
  }
  const Supernodes& SS = plan_->sn;
  const size_t tf = uf_ == kF64 ? sizeof(double) : sizeof(float);
  std::vector<char> panel(static_cast<size_t>(plan_->panel_entries) * tf, 0);
  for (int32_t s = 0; s < SS.ns; ++s) {
    const int32_t f = SS.first[s], l = SS.first[s + 1];
    const int64_t nc = l - f;
    for (int32_t j = f; j < l; ++j) {
      const int64_t cj = j - f;
      const int64_t len = Lp[j + 1] - Lp[j];
      for (int64_t t = 0; t < len; ++t) {
        const int64_t r = cj + 1 + t;  // dense front row: the pattern fills the supernode
        std::memcpy(panel.data() + (SS.panel_off[s] + panel_l_index(nc, SS.nrow[s], r, cj)) * tf,
                    static_cast<const char*>(Lx) + (Lp[j] + t) * tf, tf);
      }
    }
  }
  MPIR_HIP(hipMemcpyAsync(d_panel_.get(), panel.data(), panel.size(), hipMemcpyHostToDevice, stream_));
  MPIR_HIP(hipMemcpyAsync(d_Dinv_.get(), Dinv, n * tf, hipMemcpyHostToDevice, stream_));
  // the inverses of the diagonal blocks of big fronts, from L and D = 1/Dinv
  MPIR_CHECK(uf_ == kF64 || uf_ == kF32, "a loaded factor must be f64 or f32");
  if (uf_ == kF64) {
    reciprocal<double>(n, d_Dinv_.as<double>(), d_D_.as<double>(), stream_);
    invert_diag_blocks<double>(factor_args<double>(), dfp_, stream_);
  } else {
    reciprocal<float>(n, d_Dinv_.as<float>(), d_D_.as<float>(), stream_);
    invert_diag_blocks<float>(factor_args<float>(), dfp_, stream_);
  }
  MPIR_HIP(hipStreamSynchronize(stream_));
  analyzed_ = factored_ = true;
  precond_only_ = true;
}

void GmresLdl::round_factor16() {
  if (uf_ == kF16)
    round_panel<__half>(plan_->panel_entries, d_panel_.as<float>(), d_panel16_.as<__half>(), d_nonfinite_.as<int>(),
                        stream_);
  else if (uf_ == kBF16)
    round_panel<__hip_bfloat16>(plan_->panel_entries, d_panel_.as<float>(), d_panel16_.as<__hip_bfloat16>(),
                                d_nonfinite_.as<int>(), stream_);
}

void GmresLdl::factor(const double* Ax, bool device_ptr) {
  if (!analyzed_) throw Error(kErrState, "factor called before analyze");
  if (batch_) throw Error(kErrState, "handle holds a batch of systems: use the batch entry points");
  const double t0 = now_ms();
  factored_ = false;
  if (nnz_ > 0)
    MPIR_HIP(hipMemcpyAsync(d_ax_.get(), Ax, nnz_ * sizeof(double),
                            device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  factor_loaded(t0);
}

void GmresLdl::factor_loaded(double t0) {
  hipStream_t st = stream_;
  const int64_t ncsr = 2 * nnz_ - n_;
  // Ax_ = UW(Ax), AD_ in UW, Ax_ *= AD_ AD_ in UW (gmres_ir.hpp:96-99, 331-363)
  round_to(uw_, d_ax_.as<double>(), nnz_, st);
  if (scaling_ == kScaleRuiz) {
    // Ruiz: at most kRuizSweeps sweeps, stopping once every row's max-norm is
    // within kRuizTol of one (the tolerance is checked before each update)
    hipLaunchKernelGGL(k_fill, dim3(nblk(n_)), dim3(256), 0, st, n_, 1.0, d_ad_.as<double>());
    hipLaunchKernelGGL(k_gather, dim3(nblk(ncsr)), dim3(256), 0, st, ncsr, d_ax_.as<double>(),
                       d_csr_src_.as<int32_t>(), d_csr_x_.as<double>());
    stats_.ruiz_sweeps = 0;
    for (int sweep = 0; sweep <= kRuizSweeps; ++sweep) {
      hipLaunchKernelGGL(k_ruiz_rowmax, dim3(nblk(n_)), dim3(256), 0, st, n_, d_csr_p_.as<int32_t>(),
                         d_csr_j_.as<int32_t>(), d_csr_x_.as<double>(), d_perm_.as<int32_t>(), d_ad_.as<double>(),
                         d_r_.as<double>(), d_w_.as<double>());
      if (infnrm(d_w_.as<double>()) <= kRuizTol || sweep == kRuizSweeps) break;
      hipLaunchKernelGGL(k_ruiz_update, dim3(nblk(n_)), dim3(256), 0, st, n_, d_perm_.as<int32_t>(),
                         d_r_.as<double>(), d_ad_.as<double>());
      stats_.ruiz_sweeps = sweep + 1;
    }
  } else {
    // |a_ii| > eps(UF) selects the diagonals that scale (gmres_ir.hpp:347-363);
    // +inf disables scaling
    const double eps = scaling_ == kScaleNone ? std::numeric_limits<double>::infinity()
                       : uf_ == kF64         ? std::numeric_limits<double>::epsilon()
                       : uf_ == kF32         ? static_cast<double>(std::numeric_limits<float>::epsilon())
                       : uf_ == kF16         ? 0.0009765625    // 2^-10
                                             : 0.0078125;      // 2^-7
    hipLaunchKernelGGL(k_compute_ad, dim3(nblk(n_)), dim3(256), 0, st, n_, d_ax_.as<double>(),
                       d_diag_pos_.as<int32_t>(), eps, d_ad_.as<double>());
  }
  round_to(uw_, d_ad_.as<double>(), n_, st);
  hipLaunchKernelGGL(k_scale_entries, dim3(nblk(nnz_)), dim3(256), 0, st, nnz_, d_ax_.as<double>(),
                     d_entry_row_.as<int32_t>(), d_entry_col_.as<int32_t>(), d_ad_.as<double>(), d_axs_.as<double>());
  round_to(uw_, d_axs_.as<double>(), nnz_, st);
  hipLaunchKernelGGL(k_gather, dim3(nblk(ncsr)), dim3(256), 0, st, ncsr, d_axs_.as<double>(), d_csr_src_.as<int32_t>(),
                     d_csr_x_.as<double>());
  MPIR_HIP(hipGetLastError());
  int err = 0, pos = 0;
  if (precond_ == kPrecIlu) {
    MPIR_HIP(hipMemsetAsync(d_err_.get(), 0, sizeof(int), st));
    factor_ilu();
  } else if (uf_ == kF64) {
    const FactorArgs<double> a = factor_args<double>();
    factor_numeric<double>(a, dfp_, d_axs_.as<double>(), nnz_, 0, st);
    count_positive<double>(n_, d_D_.as<double>(), d_Dinv_.as<double>(), d_cnt_.as<int>(), d_nonfinite_.as<int>(), st);
  } else {
    const FactorArgs<float> a = factor_args<float>();
    factor_numeric<float>(a, dfp_, d_axs_.as<double>(), nnz_, rounding(), st);
    count_positive<float>(n_, d_D_.as<float>(), d_Dinv_.as<float>(), d_cnt_.as<int>(), d_nonfinite_.as<int>(), st);
    round_factor16();  // adds the 16-bit entries outside the format's range
  }
  MPIR_HIP(hipMemcpyAsync(&h_pinned_[0], d_err_.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  MPIR_HIP(hipMemcpyAsync(&h_pinned_[1], d_cnt_.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  MPIR_HIP(hipMemcpyAsync(&h_pinned_[2], d_nonfinite_.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  int nonfinite = 0;
  if (batch_) batch_count_pivots();
  MPIR_HIP(hipStreamSynchronize(st));
  std::memcpy(&err, &h_pinned_[0], sizeof(int));
  std::memcpy(&pos, &h_pinned_[1], sizeof(int));
  std::memcpy(&nonfinite, &h_pinned_[2], sizeof(int));
  stats_.positive_pivots = pos;
  stats_.factor_ms = now_ms() - t0;
  // bit 2: a dependency wait of the factorization expired (its factor is
  // then incomplete); bit 1: an exact zero pivot
  if (err & 2) throw Error(kErrInternal, "a dependency wait of the LDL^T factorization expired");
  if (err & 4) throw Error(kErrInternal, "a panel step found no slot for its diagonal block's inverse");
  if (batch_) batch_collect_pivots((err & 1) != 0);
  if (err & 1) throw Error(kErrNumeric, "zero pivot in the LDL^T factorization");
  // QDLDL_factor has no such check (qdldl.hpp:331 tests for an exact zero
  // only): a non-finite factor would turn every later solve into NaN
  if (nonfinite && (uf_ == kF16 || uf_ == kBF16) && precond_ != kPrecIlu)
    throw Error(kErrNumeric, std::to_string(nonfinite) + " factor entries overflow the " + prec_name(uf_) + " range");
  if (nonfinite)
    throw Error(kErrNumeric, std::to_string(nonfinite) + " pivots of the " + prec_name(uf_) +
                                 " factorization are not finite (overflow or NaN in the matrix or the factor)");
  factored_ = true;
}

double GmresLdl::dnrm2(const double* v) {
  hipLaunchKernelGGL(k_reduce_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, stream_, n_, v, nullptr, 1,
                     d_red_.as<double>(), nullptr);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream_, d_red_.as<double>(), kRedBlocks, 1,
                     d_small_.as<double>() + kSlotNorm, nullptr, 0, h_mirror_, nullptr);
  wait_stream();
  return h_pinned_[0];
}

double GmresLdl::infnrm(const double* v) {
  hipLaunchKernelGGL(k_reduce_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, stream_, n_, v, nullptr, 2,
                     d_red_.as<double>(), nullptr);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream_, d_red_.as<double>(), kRedBlocks, 2,
                     d_small_.as<double>() + kSlotNorm, nullptr, 0, h_mirror_, nullptr);
  wait_stream();
  return h_pinned_[0];
}

// M^{-1} applied to R vectors of n entries: each tile of the factor is read
// once for up to 8 of them
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
}

void GmresLdl::precond_apply_dev(double* x) {
  stats_.precond_applications++;
  if (precond_ == kPrecIlu) {
    if (uf_ == kF64)
      ilu_solve<double>(ilu_g_, ilu_lv_, d_ilu_l_.as<double>(), d_ilu_u_.as<double>(), d_ilu_d_.as<double>(), x, stream_);
    else
      ilu_solve<float>(ilu_g_, ilu_lv_, d_ilu_l_.as<float>(), d_ilu_u_.as<float>(), d_ilu_d_.as<float>(), x, stream_);
    return;
  }
  double* u = d_uvec_.as<double>();
  if (uf_ == kF64) ldl_solve<double, double, double>(solve_args<double, double>(d_panel_.get()), dsp_, x, u, stream_);
  else if (uf_ == kF32) ldl_solve<float, float, double>(solve_args<float, float>(d_panel_.get()), dsp_, x, u, stream_);
  else if (uf_ == kF16) ldl_solve<__half, float, double>(solve_args<__half, float>(d_panel16_.get()), dsp_, x, u, stream_);
  else ldl_solve<__hip_bfloat16, float, double>(solve_args<__hip_bfloat16, float>(d_panel16_.get()), dsp_, x, u, stream_);
}

void GmresLdl::sync_check() {
  MPIR_HIP(hipStreamSynchronize(stream_));
  if (!d_serr_.get()) return;
  int e = 0;
  MPIR_HIP(hipMemcpy(&e, d_serr_.get(), sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    MPIR_HIP(hipMemset(d_serr_.get(), 0, sizeof(int)));
    throw Error(kErrInternal, "a triangular-solve dependency wait expired");
  }
}

void GmresLdl::spmv_dev(const double* x, double* y) {
  CsrDev A{n_, d_csr_p_.as<int32_t>(), d_csr_j_.as<int32_t>(), d_csr_x_.as<double>()};
  spmv_f64(A, x, y, stream_);
}

GmDev GmresLdl::gm_dev(int m) {
  d_gm_.ensure(static_cast<size_t>(kGsHead + (m + 1) + 3 * m + static_cast<int64_t>(m + 1) * std::max(m, 1)) *
               sizeof(double));
  d_gmi_.ensure(kGiCount * sizeof(int));
  d_ir_.ensure(kIrCount * sizeof(double));
  return GmDev{d_gm_.as<double>(), d_gmi_.as<int>(), m};
}

// op 1: 2-norm, op 2: inf-norm of v into *out (device), skipped when *gate == 0
void GmresLdl::norm_to(const double* v, int op, double* out, const int* gate) {
  hipLaunchKernelGGL(k_reduce_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, stream_, n_, v, nullptr, op,
                     d_red_.as<double>(), gate);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream_, d_red_.as<double>(), kRedBlocks, op, out,
                     nullptr, 0, nullptr, gate);
}

// x := M^-1 x unless *gate == 0 (gate may be null)
void GmresLdl::precond_apply_gated(double* x, const int* gate) {
  if (precond_ == kPrecIlu) {
    IluDev g = ilu_g_;
    g.gate = gate;
    if (uf_ == kF64)
      ilu_solve<double>(g, ilu_lv_, d_ilu_l_.as<double>(), d_ilu_u_.as<double>(), d_ilu_d_.as<double>(), x, stream_);
    else
      ilu_solve<float>(g, ilu_lv_, d_ilu_l_.as<float>(), d_ilu_u_.as<float>(), d_ilu_d_.as<float>(), x, stream_);
    return;
  }
  SolveDev d = dsp_.d;
  d.gate = gate;
  double* u = d_uvec_.as<double>();
  if (uf_ == kF64) ldl_solve_on<double, double, double>(solve_args<double, double>(d_panel_.get()), dsp_, d, x, u, stream_);
  else if (uf_ == kF32) ldl_solve_on<float, float, double>(solve_args<float, float>(d_panel_.get()), dsp_, d, x, u, stream_);
  else if (uf_ == kF16)
    ldl_solve_on<__half, float, double>(solve_args<__half, float>(d_panel16_.get()), dsp_, d, x, u, stream_);
  else
    ldl_solve_on<__hip_bfloat16, float, double>(solve_args<__hip_bfloat16, float>(d_panel16_.get()), dsp_, d, x, u,
                                                stream_);
}

// Waits until the device has published record `target` in mirror word
// `slot`: polls the mapped word (no HIP call), yields the core after 200 us,
// and every millisecond asks the stream whether it failed or went idle.
void GmresLdl::wait_mirror(int slot, double target) {
  volatile double* mv = h_pinned_;
  const auto t0 = std::chrono::steady_clock::now();
  auto last = t0;
  for (;;) {
    if (mv[slot] >= target) return;
    const auto t = std::chrono::steady_clock::now();
    if (t - t0 > std::chrono::microseconds(200)) std::this_thread::yield();
    if (t - last > std::chrono::milliseconds(1)) {
      last = t;
      const hipError_t e = hipStreamQuery(stream_);
      if (e == hipSuccess) {
        if (mv[slot] >= target) return;
        throw Error(kErrInternal, "the device did not publish its iteration record");
      }
      if (e != hipErrorNotReady) MPIR_HIP(e);
    }
  }
}

// PrecondGmres (gmres_ir.hpp:636-741) on the device: res = V y from the
// Krylov basis of M^-1 K_s started at M^-1 b. The call is skipped when
// *gate == 0 (res = 0). The host queues Arnoldi step k + 1 while step k
// runs and stops queueing once a finished step reports the iteration over,
// so at most one queued step finds its gate closed.
void GmresLdl::precond_gmres(const double* b, double* res, const int* gate) {
  hipStream_t st = stream_;
  const int n = n_;
  max_gmres_ = std::min<int64_t>(max_gmres_, n);  // persists, as gmres_ir.hpp:643
  const int m = static_cast<int>(max_gmres_);
  MPIR_CHECK(m + 2 <= kMirStep - 8, "GMRES iteration limit above 3990");
  const int64_t ldv = n;
  d_V_.ensure(static_cast<size_t>(m + 1) * ldv * sizeof(double));
  d_hcol_.ensure(static_cast<size_t>(m + 2) * sizeof(double));
  d_red_.ensure(static_cast<size_t>(std::max(m + 1, 2)) * kRedBlocks * sizeof(double));
  d_small_.ensure(static_cast<size_t>(kSlotH + m + 2) * sizeof(double));
  const GmDev gd = gm_dev(m);
  double* V = d_V_.as<double>();
  const int* active = gd.gi + kGiActive;
  norm_to(b, 1, gd.gs + kGsRhoB, gate);
  hipLaunchKernelGGL(k_gm_begin, dim3(1), dim3(1), 0, st, gd, gate, tol_);
  hipLaunchKernelGGL(k_copy_gated, dim3(nblk(n)), dim3(256), 0, st, n, b, V, active);
  precond_apply_gated(V, active);
  norm_to(V, 1, gd.gs + kGsRho, active);
  hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(1), 0, st, gd, tol_);
  hipLaunchKernelGGL(k_scale_dev, dim3(nblk(n)), dim3(256), 0, st, n, V, gd.gs + kGsInv, 0, V, active);
  const unsigned red_blocks = static_cast<unsigned>(std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + 255) / 256)));
  double* part = d_red_.as<double>();
  double* h1 = d_hcol_.as<double>();
  double* h2 = d_small_.as<double>() + kSlotH;
  CsrDev A{n_, d_csr_p_.as<int32_t>(), d_csr_j_.as<int32_t>(), d_csr_x_.as<double>()};
  A.gate = active;
  for (int k = 0; k < m; ++k) {
    double* w = V + ldv * (k + 1);
    spmv_f64(A, V + ldv * k, w, st);
    precond_apply_gated(w, active);
    // CGS2: two passes of h = V^T w, w -= V h
    const int k1 = k + 1;
    for (int pass = 0; pass < 2; ++pass) {
      double* hp = pass == 0 ? h1 : h2;
      hipLaunchKernelGGL(k_mdot_partial, dim3(red_blocks), dim3(kRedThreads), 0, st, n, V, ldv, k1, w, part, active);
      hipLaunchKernelGGL(k_mdot_final, dim3(k1), dim3(kRedThreads), 0, st, part, static_cast<int>(red_blocks), hp, 0,
                         active);
      hipLaunchKernelGGL(k_sub_basis, dim3(nblk(n)), dim3(256), 0, st, n, V, ldv, k1, hp, w, active);
    }
    norm_to(w, 1, gd.gs + kGsHn, active);
    const double ord = static_cast<double>(++step_seq_);
    hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(1), 0, st, gd, k, tol_, h1, h2, h_mirror_, ord);
    hipLaunchKernelGGL(k_scale_dev, dim3(nblk(n)), dim3(256), 0, st, n, w, gd.gs + kGsInv, 0, w, active);
    MPIR_HIP(hipGetLastError());
    if (k >= 1) {  // step k is queued; see whether step k - 1 ended the iteration
      wait_mirror(kMirStep, ord - 1);
      if (static_cast<volatile double*>(h_pinned_)[kMirActive] == 0.0) break;
    }
  }
  hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(1), 0, st, gd);
  hipLaunchKernelGGL(k_gm_combine, dim3(nblk(n)), dim3(256), 0, st, n, V, ldv, gd.gi + kGiK, gd.y(), res);
}

// GmresLDLIR::Solve (gmres_ir.hpp:534-634) on the device. The host queues
// each refinement iteration and waits once per iteration for its record
// (stop reason, residual, correction norm, counts); the norms, the scaling
// and the stopping rules are computed on the device.
void GmresLdl::solve_one(const double* b, double* x) {
  hipStream_t st = stream_;
  const int n = n_;
  CsrDev A{n_, d_csr_p_.as<int32_t>(), d_csr_j_.as<int32_t>(), d_csr_x_.as<double>()};
  double* r = d_r_.as<double>();
  double* d = d_d_.as<double>();
  double* tmp = d_w_.as<double>();
  const double t_start = now_ms();
  // eps of the working precision in the small-update test (gmres_ir.hpp:599)
  const double eps_uw = uw_ == kF32 ? static_cast<double>(std::numeric_limits<float>::epsilon())
                                    : std::numeric_limits<double>::epsilon();
  const GmDev gd = gm_dev(static_cast<int>(std::min<int64_t>(max_gmres_, n)));
  double* ir = d_ir_.as<double>();
  const int* irgate = gd.gi + kGiIrGate;
  const int32_t base_k = stats_.gmres_iterations, base_apps = stats_.precond_applications;
  volatile double* mv = h_pinned_;
  hipLaunchKernelGGL(k_ir_reset, dim3(1), dim3(1), 0, st, ir, gd.gi);
  norm_to(b, 2, ir + kIrBnorm, nullptr);
  precond_gmres(b, x, nullptr);
  round_to(uw_, x, n, st);
  residual_f64(A, x, b, r, st);
  round_to(ur_, r, n, st);
  norm_to(r, 2, ir + kIrR0, nullptr);
  CsrDev Ag = A;
  Ag.gate = irgate;
  stats_.stop_reason = 0;
  log_ = SolveLog();
  double r_norm = 0.0;
  bool first = true;
  // A non-finite right-hand side, iterate or residual ends the solve with
  // MPIR_ERR_NUMERIC. The reference has no such check: its InfNrm skips a
  // NaN that is not the first entry (ops.hpp:324-329), and the loop would go
  // on with NaN.
  auto take_record = [&](int64_t it) {
    const double sq = static_cast<double>(++ir_seq_);
    hipLaunchKernelGGL(k_ir_check, dim3(1), dim3(1), 0, st, ir, gd.gi, static_cast<int>(it), tol_, eps_uw, h_mirror_,
                       sq);
    MPIR_HIP(hipGetLastError());
    wait_mirror(kMirIrSeq, sq);
    stats_.gmres_iterations = base_k + static_cast<int32_t>(mv[kMirIr + 4]);
    stats_.precond_applications = base_apps + static_cast<int32_t>(mv[kMirIr + 5]);
    if (first) {
      if (!std::isfinite(mv[kMirIr + 6])) throw Error(kErrNumeric, "the right-hand side is not finite");
      if (!std::isfinite(mv[kMirIr + 7])) throw Error(kErrNumeric, "the first iterate is not finite");
      log_.initial_residual = mv[kMirIr + 7];
      first = false;
    }
  };
  int64_t iter;
  for (iter = 1; iter <= max_ir_; ++iter) {
    norm_to(r, 2, ir + kIrScale, nullptr);
    // an exactly zero residual would make the reference scale by 1/0
    // (gmres_ir.hpp:562-563) and carry NaN forward; x is exact: reason 3
    hipLaunchKernelGGL(k_ir_gate, dim3(1), dim3(1), 0, st, ir, gd.gi);
    hipLaunchKernelGGL(k_scale_dev, dim3(nblk(n)), dim3(256), 0, st, n, r, ir + kIrScale, 1, tmp, irgate);
    round_to(uw_, tmp, n, st);
    const double t_inner = now_ms();
    precond_gmres(tmp, d, irgate);
    hipLaunchKernelGGL(k_scale_dev, dim3(nblk(n)), dim3(256), 0, st, n, d, ir + kIrScale, 0, d, irgate);
    round_to(uw_, d, n, st);
    norm_to(d, 2, ir + kIrDnorm, irgate);
    norm_to(x, 2, ir + kIrXnorm, irgate);  // ||x|| before the update (gmres_ir.hpp:575-577)
    hipLaunchKernelGGL(k_add_gated, dim3(nblk(n)), dim3(256), 0, st, n, x, d, irgate);
    round_to(uw_, x, n, st);
    residual_f64(Ag, x, b, r, st);
    round_to(ur_, r, n, st);
    norm_to(r, 2, ir + kIrRraw, irgate);
    take_record(iter);
    const double inner_ms = now_ms() - t_inner;  // queueing and completion of this iteration
    stats_.ir_iterations = static_cast<int32_t>(iter);
    const int stop = static_cast<int>(mv[kMirIr + 1]);
    if (stop == 3) {
      stats_.stop_reason = 3;
      r_norm = 0.0;
      break;
    }
    r_norm = mv[kMirIr + 2];
    const double d_norm = mv[kMirIr + 3];
    if (!std::isfinite(r_norm) || !std::isfinite(d_norm))
      throw Error(kErrNumeric, "the residual is not finite at refinement iteration " + std::to_string(iter));
    // the VERBOSE log's record of the iteration (gmres_ir.hpp:583-596)
    log_.iterations.push_back({static_cast<int32_t>(iter), r_norm, d_norm, now_ms() - t_start, inner_ms});
    if (stop != 0) {
      stats_.stop_reason = stop;
      break;
    }
  }
  if (max_ir_ == 0) {
    take_record(0);
    r_norm = mv[kMirIr + 2];
  }
  stats_.final_rel_residual = r_norm;
  log_.stop_reason = stats_.stop_reason;
  log_.total_iterations = static_cast<int32_t>(std::min<int64_t>(iter, max_ir_));
  log_.total_time_ms = now_ms() - t_start;
}

// The solver log in the schema of the reference's VERBOSE build
// (gmres_ir.hpp:119-124, 305-314, 539-631): max iterations, tolerance and one
// "gmres_ir" event with its per-iteration records for the last solve.
std::string GmresLdl::log_json() const {
  auto num = [](double v) {
    if (!std::isfinite(v)) return std::string("null");
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return std::string(buf);
  };
  static const char* reasons[] = {"max_iterations_reached", "converged_small_update",
                                  "converged_residual_stagnation", "converged_zero_residual"};
  static const char* details[] = {"", "d_norm <= x_norm * (10*eps)", "abs(r_norm - r_norm_prev) < tol",
                                  "residual exactly zero"};
  std::string o = "{\n  \"max_outer_iterations\": " + std::to_string(max_ir_) +
                  ",\n  \"max_inner_iterations\": " + std::to_string(max_gmres_) + ",\n  \"tolerance\": " +
                  num(tol_) + ",\n  \"events\": [\n    {\n      \"name\": \"gmres_ir\",\n      \"initial_residual\": " +
                  num(log_.initial_residual) + ",\n      \"iterations\": [";
  for (size_t i = 0; i < log_.iterations.size(); ++i) {
    const auto& e = log_.iterations[i];
    o += std::string(i ? "," : "") + "\n        {\"iteration\": " + std::to_string(e.iteration) +
         ", \"residual\": " + num(e.residual) + ", \"d_norm\": " + num(e.d_norm) +
         ", \"timestamp_ms\": " + num(e.timestamp_ms) + ", \"inner_time_ms\": " + num(e.inner_time_ms) + "}";
  }
  const int sr = std::max(0, std::min(3, log_.stop_reason));
  o += std::string(log_.iterations.empty() ? "]" : "\n      ]") + ",\n      \"stop_reason\": \"" + reasons[sr] + "\"";
  if (sr) o += ",\n      \"stop_reason_detail\": \"" + std::string(details[sr]) + "\"";
  o += ",\n      \"total_iterations\": " + std::to_string(log_.total_iterations) + ",\n      \"total_time_ms\": " +
       num(log_.total_time_ms) + "\n    }\n  ]\n}\n";
  return o;
}

void GmresLdl::solve(const double* b, double* x, int nrhs, int64_t ld, bool device_ptr) {
  if (batch_) throw Error(kErrState, "handle holds a batch of systems: use the batch entry points");
  if (!factored_) throw Error(kErrState, "solve called before a successful factorization");
  if (precond_only_) throw Error(kErrState, "handle holds a loaded factor only; call compute first");
  MPIR_CHECK(nrhs >= 1 && ld >= n_, "bad right-hand-side layout");
  const double t0 = now_ms();
  hipStream_t st = stream_;
  const int n = n_;
  stats_.ir_iterations = 0;
  stats_.gmres_iterations = 0;
  stats_.precond_applications = 0;
  for (int c = 0; c < nrhs; ++c) {
    const double* bc = b + ld * c;
    double* xc = x + ld * c;
    const double* bin;
    if (device_ptr) {
      bin = bc;
    } else {
      MPIR_HIP(hipMemcpyAsync(d_r_.get(), bc, n * sizeof(double), hipMemcpyHostToDevice, st));
      bin = d_r_.as<double>();
    }
    hipLaunchKernelGGL(k_perm_in, dim3(nblk(n)), dim3(256), 0, st, n, bin, d_perm_.as<int32_t>(), d_ad_.as<double>(),
                       d_b_.as<double>());
    round_to(uw_, d_b_.as<double>(), n, st);  // b_scaled = VectorMultiply<UW>(b, AD_)
    solve_one(d_b_.as<double>(), d_x_.as<double>());
    double* xout = device_ptr ? xc : d_d_.as<double>();
    hipLaunchKernelGGL(k_perm_out, dim3(nblk(n)), dim3(256), 0, st, n, d_x_.as<double>(), d_perm_.as<int32_t>(),
                       d_ad_.as<double>(), xout);
    round_to(uw_, xout, n, st);  // VectorMultiply<UW>(x, AD_)
    if (!device_ptr) MPIR_HIP(hipMemcpyAsync(xc, xout, n * sizeof(double), hipMemcpyDeviceToHost, st));
    sync_check();
  }
  stats_.solve_ms = now_ms() - t0;
}

int64_t GmresLdl::trace_factor(uint64_t* out, int64_t cap) {
  MPIR_CHECK(analyzed_ && !precond_only_ && precond_ == kPrecLdl, "trace_factor needs an analyzed L D L^T handle with values");
  const int64_t nt = dfp_.dag.ntask;
  if (cap == 0 || nt == 0) return nt;
  hipStream_t st = stream_;
  d_ftrace_.ensure(static_cast<size_t>(nt) * 4 * sizeof(uint64_t));
  MPIR_HIP(hipMemsetAsync(d_ftrace_.get(), 0, static_cast<size_t>(nt) * 4 * sizeof(uint64_t), st));
  dfp_.dag.trace = d_ftrace_.as<unsigned long long>();
  try {
    if (uf_ == kF64) factor_numeric<double>(factor_args<double>(), dfp_, d_axs_.as<double>(), nnz_, 0, st);
    else factor_numeric<float>(factor_args<float>(), dfp_, d_axs_.as<double>(), nnz_, rounding(), st);
  } catch (...) {
    dfp_.dag.trace = nullptr;
    throw;
  }
  dfp_.dag.trace = nullptr;
  const int64_t n = std::min(nt, cap / 4) * 4;
  MPIR_HIP(hipMemcpyAsync(out, d_ftrace_.get(), static_cast<size_t>(n) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  MPIR_HIP(hipStreamSynchronize(st));
  return nt;
}

int32_t GmresLdl::supernode_info(int32_t* level, int32_t* ncols, int32_t* nrows, int32_t cap) const {
  MPIR_CHECK(analyzed_ && plan_ != nullptr && precond_ == kPrecLdl, "supernode_info needs an analyzed L D L^T handle");
  const Supernodes& S = plan_->sn;
  if (cap >= S.ns) {
    for (int32_t s = 0; s < S.ns; ++s) {
      level[s] = S.level[s];
      ncols[s] = S.first[s + 1] - S.first[s];
      nrows[s] = S.nrow[s];
    }
  }
  return S.ns;
}

int64_t GmresLdl::trace_precond(uint64_t* out, int64_t cap, int64_t* nfwd) {
  MPIR_CHECK(factored_ && precond_ == kPrecLdl, "trace_precond needs a factored L D L^T handle");
  MPIR_CHECK(dsp_.schedule == kScheduleDag, "trace_precond traces the dependency-driven schedule");
  const int64_t nt = static_cast<int64_t>(dsp_.d.nfwd) + dsp_.d.nbwd;
  *nfwd = dsp_.d.nfwd;
  if (cap == 0 || nt == 0) return nt;
  hipStream_t st = stream_;
  d_strace_.ensure(static_cast<size_t>(nt) * 4 * sizeof(uint64_t));
  MPIR_HIP(hipMemsetAsync(d_strace_.get(), 0, static_cast<size_t>(nt) * 4 * sizeof(uint64_t), st));
  double* x = d_w_.as<double>();
  MPIR_HIP(hipMemcpyAsync(x, d_ad_.get(), n_ * sizeof(double), hipMemcpyDeviceToDevice, st));
  dsp_.d.trace = d_strace_.as<unsigned long long>();
  try {
    precond_apply_dev(x);
  } catch (...) {
    dsp_.d.trace = nullptr;
    throw;
  }
  dsp_.d.trace = nullptr;
  const int64_t n = std::min(nt, cap / 4) * 4;
  MPIR_HIP(hipMemcpyAsync(out, d_strace_.get(), static_cast<size_t>(n) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  sync_check();
  return nt;
}

double GmresLdl::time_kernels(int what, int reps) {
  if (!factored_) throw Error(kErrState, "time_kernels needs a factorization");
  MPIR_CHECK(reps >= 1 && what >= 0 && what <= 3, "bad timing request");
  if (what == 3) return time_precond_slots(std::min(8, std::max(1, b_slots_)), reps);
  if (what == 2) MPIR_CHECK(!precond_only_, "no values to refactor");
  hipStream_t st = stream_;
  double* x = d_w_.as<double>();
  double* y = d_d_.as<double>();
  // a deterministic nonzero input: the scaling vector
  MPIR_HIP(hipMemcpyAsync(x, d_ad_.get(), n_ * sizeof(double), hipMemcpyDeviceToDevice, st));
  hipEvent_t e0, e1;
  MPIR_HIP(hipEventCreate(&e0));
  MPIR_HIP(hipEventCreate(&e1));
  auto run = [&]() {
    if (what == 0) {
      spmv_dev(x, y);
    } else if (what == 1) {
      precond_apply_dev(x);
    } else if (precond_ == kPrecIlu) {
      factor_ilu();
    } else if (uf_ == kF64) {
      const FactorArgs<double> a = factor_args<double>();
      factor_numeric<double>(a, dfp_, d_axs_.as<double>(), nnz_, 0, st);
    } else {
      const FactorArgs<float> a = factor_args<float>();
      factor_numeric<float>(a, dfp_, d_axs_.as<double>(), nnz_, rounding(), st);
      round_factor16();
    }
  };
  run();  // warm
  MPIR_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) run();
  MPIR_HIP(hipEventRecord(e1, st));
  MPIR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPIR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  sync_check();
  return static_cast<double>(ms) / reps;
}

double GmresLdl::time_precond_slots(int rhs, int reps) {
  hipStream_t st = stream_;
  d_V_.ensure(static_cast<size_t>(rhs) * n_ * sizeof(double));
  double* x = d_V_.as<double>();
  // a deterministic nonzero input: the scaling vector in every slot
  for (int r = 0; r < rhs; ++r)
    MPIR_HIP(hipMemcpyAsync(x + static_cast<int64_t>(r) * n_, d_ad_.get(), n_ * sizeof(double),
                            hipMemcpyDeviceToDevice, st));
  const int64_t apps = stats_.precond_applications;
  precond_apply_slots(x, rhs);  // warm
  hipEvent_t e0, e1;
  MPIR_HIP(hipEventCreate(&e0));
  MPIR_HIP(hipEventCreate(&e1));
  MPIR_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) precond_apply_slots(x, rhs);
  MPIR_HIP(hipEventRecord(e1, st));
  MPIR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPIR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  stats_.precond_applications = apps;
  sync_check();
  return static_cast<double>(ms) / reps;
}

void GmresLdl::apply_tuning() {
  FactorPlan& P = *plan_;
  if (!same_tuning(P.ftuning, ftune_)) build_factor_steps(P, ftune_);
  if (!same_tuning(P.tuning, stune_)) build_solve_schedule(P, stune_);
  if (factor_schedule_ == kScheduleDag && !P.dag_pool) build_update_pool(P, true);
}

void GmresLdl::set_factor_schedule(int sched) {
  MPIR_CHECK(sched == kScheduleLevels || sched == kScheduleDag, "factor schedule must be 0 (levels) or 1 (dag)");
  factor_schedule_ = sched;
  dfp_.use_dag = sched == kScheduleDag;
  if (sched == kScheduleDag && plan_ && analyzed_ && !plan_->dag_pool) {
    // the analyzed plan's update pool reuses blocks across levels: lay it
    // out again for the dependency-driven launch (the factor must be
    // computed again)
    build_update_pool(*plan_, true);
    upload_vec(d_upd_off_, plan_->upd_off, stream_);
    const size_t tf = uf_ == kF64 ? sizeof(double) : sizeof(float);
    d_upool_.alloc((std::max<int64_t>(1, plan_->upd_pool) + kPanelSlack) * tf);
    upload_factor_dag();
    factored_ = false;
    MPIR_HIP(hipStreamSynchronize(stream_));
  }
}

void GmresLdl::set_tuning(const FactorTuning& f, const SolveTuning& s) {
  MPIR_CHECK(f.upd_group >= 1 && f.upd_group <= 64 && f.super_tiles >= 1 && f.super_tiles <= 64 && s.narrow_rows >= 0,
             "tuning out of range");
  ftune_ = f;
  stune_ = s;
  if (plan_ && analyzed_) {
    const int slots = std::max(1, solve_slots_);
    apply_tuning();
    upload_plan();
    set_solve_slots(slots);
  }
}

AutotuneResult GmresLdl::autotune(int reps) {
  if (!factored_) throw Error(kErrState, "autotune needs a factorization");
  MPIR_CHECK(precond_ == kPrecLdl && !precond_only_, "autotune needs an L D L^T factor of loaded values");
  MPIR_CHECK(reps >= 1 && reps <= 100, "reps must be in 1 .. 100");
  FactorPlan& P = *plan_;
  AutotuneResult res;
  const int slots = std::max(1, solve_slots_);
  auto install = [&]() {
    upload_plan();
    set_solve_slots(slots);
  };
  // the factor's step schedule: update grouping x super-block edge, the
  // tuning in place first
  std::vector<FactorTuning> fc{P.ftuning};
  for (int g : {2, 4, 8})
    for (int t : {4, 8, 16}) {
      const FactorTuning c{g, t};
      if (!same_tuning(c, P.ftuning)) fc.push_back(c);
    }
  double best = 0.0;
  FactorTuning fbest = P.ftuning;
  for (size_t i = 0; i < fc.size(); ++i) {
    if (i > 0) {
      build_factor_steps(P, fc[i]);
      install();
    }
    const double ms = time_kernels(2, reps);
    if (i == 0) res.factor_ms_before = ms;
    if (i == 0 || ms < best) {
      best = ms;
      fbest = fc[i];
    }
  }
  res.factor = fbest;
  res.factor_ms = best;
  build_factor_steps(P, fbest);
  // the solves: narrow-front row limit x schedule
  const SolveTuning s0 = P.tuning;
  const int sched0 = schedule_;
  struct SC {
    SolveTuning t;
    int sched;
  };
  std::vector<SC> sc{{s0, sched0}};
  for (int rows : {0, 256, 512, 1024, 2048, 4096})
    for (int sched : {kScheduleDag, kScheduleLevels}) {
      const SolveTuning t{rows};
      if (!(same_tuning(t, s0) && sched == sched0)) sc.push_back({t, sched});
    }
  // M^{-1} is timed on as many right-hand sides as the handle's solves apply
  // it to at once (a batch: its last solve's, up to 8; else 1)
  const int rhs = batch_ ? std::min(8, std::max(1, b_slots_)) : 1;
  SC sbest = sc[0];
  for (size_t i = 0; i < sc.size(); ++i) {
    build_solve_schedule(P, sc[i].t);
    schedule_ = sc[i].sched;
    install();
    const double ms = rhs == 1 ? time_kernels(1, reps) : time_precond_slots(rhs, reps);
    if (i == 0) res.precond_ms_before = ms;
    if (i == 0 || ms < best) {
      best = ms;
      sbest = sc[i];
    }
  }
  res.solve = sbest.t;
  res.schedule = sbest.sched;
  res.precond_ms = best;
  res.candidates = static_cast<int32_t>(fc.size() + sc.size());
  // keep the choice, and the factor computed with it
  ftune_ = fbest;
  stune_ = sbest.t;
  schedule_ = sbest.sched;
  build_solve_schedule(P, stune_);
  install();
  time_kernels(2, 1);
  return res;
}

template <class TF>
void GmresLdl::luir_impl(const double* b, double* x, int maxiter, LuirResult* out) {
  hipStream_t st = stream_;
  const int n = n_;
  double* db = d_b_.as<double>();
  double* dx = d_x_.as<double>();
  double* dr = d_r_.as<double>();
  double* dw = d_w_.as<double>();
  TF* c = reinterpret_cast<TF*>(d_d_.get());
  TF* uf = reinterpret_cast<TF*>(d_uvec_.get());
  const SolveArgs<TF, TF> sa = solve_args<TF, TF>(d_panel_.get());
  CsrDev A{n_, d_csr_p_.as<int32_t>(), d_csr_j_.as<int32_t>(), d_csr_x_.as<double>()};
  // max |v| over a TF vector, exactly (the values are widened to double)
  auto absmax = [&](const TF* v) {
    hipLaunchKernelGGL((k_convert<TF, double>), dim3(nblk(n)), dim3(256), 0, st, n, v, dw);
    return infnrm(dw);
  };
  MPIR_HIP(hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, st));
  const double bn = infnrm(db);
  const double t0 = now_ms();
  out->steps = 0;
  if (maxiter == 0) {
    MPIR_HIP(hipMemcpyAsync(dx, db, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    ldl_solve<TF, TF, double>(sa, dsp_, dx, d_uvec_.as<double>(), st);
    MPIR_HIP(hipStreamSynchronize(st));
    out->time_solve = now_ms() - t0;
    out->time_ref = 0.0;
  } else {
    // x0 in the factor precision (luir.hpp:173-178)
    hipLaunchKernelGGL((k_convert<double, TF>), dim3(nblk(n)), dim3(256), 0, st, n, db, c);
    ldl_solve<TF, TF, TF>(sa, dsp_, c, uf, st);
    hipLaunchKernelGGL((k_convert<TF, double>), dim3(nblk(n)), dim3(256), 0, st, n, c, dx);
    MPIR_HIP(hipStreamSynchronize(st));
    out->time_solve = now_ms() - t0;
    const double t1 = now_ms();
    double d0 = std::numeric_limits<double>::max();
    for (int it = 1; it <= maxiter; ++it) {
      residual_f64(A, dx, db, dr, st);
      hipLaunchKernelGGL((k_convert<double, TF>), dim3(nblk(n)), dim3(256), 0, st, n, dr, c);
      const double nc = absmax(c);
      if (nc == 0.0) {  // x is exact (the reference would divide by zero here)
        out->steps = it;
        break;
      }
      hipLaunchKernelGGL(k_scale_by<TF>, dim3(nblk(n)), dim3(256), 0, st, n, 1. / nc, c);
      ldl_solve<TF, TF, TF>(sa, dsp_, c, uf, st);
      hipLaunchKernelGGL(k_scale_by<TF>, dim3(nblk(n)), dim3(256), 0, st, n, nc, c);
      const double dn = absmax(c);
      const double xn = infnrm(dx);
      const bool stop = dn / xn <= 10 * std::numeric_limits<double>::epsilon() || dn / d0 >= 0.5;
      out->steps = it;
      hipLaunchKernelGGL(k_add_from<TF>, dim3(nblk(n)), dim3(256), 0, st, n, c, dx);
      if (stop) break;
      d0 = dn;
    }
    MPIR_HIP(hipStreamSynchronize(st));
    out->time_ref = now_ms() - t1;
  }
  // relative residual with the residual rounded to the factor precision
  residual_f64(A, dx, db, dr, st);
  hipLaunchKernelGGL((k_convert<double, TF>), dim3(nblk(n)), dim3(256), 0, st, n, dr, c);
  out->rel_residual = absmax(c) / bn;
  MPIR_HIP(hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, st));
  sync_check();
}

void GmresLdl::luir(const double* b, double* x, int maxiter, LuirResult* out) {
  if (!factored_ || precond_only_) throw Error(kErrState, "LUIR needs a factorization of the matrix");
  if (precond_ != kPrecLdl) throw Error(kErrState, "LUIR runs on an L D L^T factorization");
  MPIR_CHECK(ordering_ == kOrderNatural && scaling_ == kScaleNone, "LUIR runs on the matrix as given");
  MPIR_CHECK(maxiter >= 0, "maxiter must be non-negative");
  if (uf_ == kF64) luir_impl<double>(b, x, maxiter, out);
  else if (uf_ == kF32) luir_impl<float>(b, x, maxiter, out);
  else throw Error(kErrInvalid, "LUIR factor precision must be f64 or f32");
}

void GmresLdl::export_factor(int64_t* Lp, int32_t* Li, void* Lx, void* Dv, void* Dinv, int prec) {
  if (batch_) throw Error(kErrState, "export_factor works on single-system handles");
  if (precond_ == kPrecIlu) throw Error(kErrState, "the handle holds FGPILU factors (export_ilu)");
  if (!factored_ && stats_.positive_pivots < 0) throw Error(kErrState, "no factor to export");
  const Supernodes& S = plan_->sn;
  const size_t tf = uf_ == kF64 ? 8 : 4;
  std::vector<char> panel(plan_->panel_entries * tf), D(n_ * tf), Di(n_ * tf);
  MPIR_HIP(hipMemcpyAsync(panel.data(), d_panel_.get(), panel.size(), hipMemcpyDeviceToHost, stream_));
  MPIR_HIP(hipMemcpyAsync(D.data(), d_D_.get(), D.size(), hipMemcpyDeviceToHost, stream_));
  MPIR_HIP(hipMemcpyAsync(Di.data(), d_Dinv_.get(), Di.size(), hipMemcpyDeviceToHost, stream_));
  MPIR_HIP(hipStreamSynchronize(stream_));
  auto get = [&](const std::vector<char>& buf, int64_t i) -> double {
    if (tf == 8) { double v; std::memcpy(&v, buf.data() + i * 8, 8); return v; }
    float v; std::memcpy(&v, buf.data() + i * 4, 4); return v;
  };
  auto put = [&](void* dst, int64_t i, double v) {
    if (prec == kF64) static_cast<double*>(dst)[i] = v;
    else if (prec == kF32) static_cast<float*>(dst)[i] = static_cast<float>(v);
    else if (prec == kF16) static_cast<__half*>(dst)[i] = __float2half(static_cast<float>(v));
    else static_cast<__hip_bfloat16*>(dst)[i] = __float2bfloat16(static_cast<float>(v));
  };
  // true pattern of each column (relaxed blocks also store explicit zeros)
  std::vector<int64_t> TLp;
  std::vector<int32_t> TLi;
  const bool use_true = relaxed_ && !Cp_.empty();
  if (use_true) l_pattern(n_, Cp_.data(), Ci_.data(), parent_.data(), Lnz_.data(), TLp, TLi);
  int64_t q = 0;
  Lp[0] = 0;
  for (int32_t s = 0; s < S.ns; ++s) {
    const int32_t f = S.first[s], l = S.first[s + 1];
    const int64_t m = (l - f) + S.nrow[s];
    const int32_t* R = S.rows.data() + S.rows_off[s];
    for (int32_t j = f; j < l; ++j) {
      const int64_t cj = j - f;
      auto row_of = [&](int64_t r) { return r < (l - f) ? static_cast<int32_t>(f + r) : R[r - (l - f)]; };
      if (use_true) {
        int64_t r = cj + 1;
        for (int64_t t = TLp[j]; t < TLp[j + 1]; ++t) {
          const int32_t row = TLi[t];
          while (r < m && row_of(r) < row) ++r;
          if (r >= m || row_of(r) != row) throw Error(kErrInternal, "factor pattern outside its supernode");
          if (Li) Li[q] = row;
          if (Lx) put(Lx, q, get(panel, S.panel_off[s] + panel_l_index(l - f, S.nrow[s], r, cj)));
          ++q;
        }
      } else {
        for (int64_t r = cj + 1; r < m; ++r) {
          if (Li) Li[q] = row_of(r);
          if (Lx) put(Lx, q, get(panel, S.panel_off[s] + panel_l_index(l - f, S.nrow[s], r, cj)));
          ++q;
        }
      }
      Lp[j + 1] = q;
      if (Dv) put(Dv, j, get(D, j));
      if (Dinv) put(Dinv, j, get(Di, j));
    }
  }
}

}  // namespace mpir
