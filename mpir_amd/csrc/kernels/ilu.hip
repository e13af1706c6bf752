// FGPILU on gfx950: the fine-grained parallel ILU of the reference's
// GmresLDLIR::Compute (include/MPIR/gmres_ir.hpp:412-518 of yex33/MPIR,
// Chow & Patel's fixed-point sweeps) and its M^{-1}, TriangularSolve
// (:743-774).
//
// The reference's sweep visits columns ascending and the entries of a column
// top to bottom, updating U and L in place. Every value an entry reads
// (L[row, k] and U[k, col] for k < row, 1 / U_kk) belongs to an earlier column
// or to an earlier entry of the same column, so that sweep leaves the exact
// ILU factors of the pattern, and a second sweep reproduces them bit for bit.
// Here the same sequence runs level by level: columns whose pattern rows
// lie in earlier levels only form a level, one thread per column walks its
// entries in order with the reference's operations (no contraction into fused
// multiply-adds, as the host code it mirrors), so the factors equal the
// reference's single-threaded ones exactly. SparseLDotU (:776-799) is the
// merge of row r of L with column c of U below the cut. The nonlinear
// residual NLResNorm (:801-824) sums |A - L U| over the pattern: per-entry
// terms in the factor precision, block sums and a final sum in double.
//
// The triangular solves run level by level (host-built level sets): L y = b
// by rows of L (unit diagonal), then U x = y by rows of U, subtracting in
// descending column order before the product with 1 / U_jj, in FP64: the
// reference's column-oriented loops, in the same order per entry.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common.hpp"
#include "kernels.hpp"

namespace mpir {

namespace {

constexpr int kIluThreads = 256;

template <class T>
__device__ __forceinline__ T ldotu(const IluDev& g, const T* Lx, const T* Ux, int row, int col, int cut) {
#pragma clang fp contract(off)
  int i1 = g.Up[row], i2 = g.Up[col];
  const int e1 = g.Up[row + 1], e2 = g.Up[col + 1];
  T sum = T(0);
  while (i1 < e1 && i2 < e2) {
    const int kL = g.Ui[i1], kU = g.Ui[i2];
    if (kL >= cut || kU >= cut) break;
    if (kL == kU) {
      sum += Lx[i1] * Ux[i2];
      ++i1;
      ++i2;
    } else if (kL < kU) {
      ++i1;
    } else {
      ++i2;
    }
  }
  return sum;
}

// initial guesses (:414-441): U = L = A on the pattern, 1 / A_jj
template <class T>
__global__ void k_ilu_init(IluDev g, const double* __restrict__ ax, T* __restrict__ av, T* __restrict__ Ux,
                           T* __restrict__ Lx, T* __restrict__ Di) {
  const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (idx >= g.nnz) return;
  const int q = g.src[idx];
  const T v = q >= 0 ? static_cast<T>(ax[q]) : T(0);
  av[idx] = v;
  Ux[idx] = v;
  Lx[idx] = v;
  if (g.Ui[idx] == g.col[idx]) Di[g.col[idx]] = T(1) / v;
}

// one sweep over the columns of one level (:464-492): per entry (row, col)
// U_rc = A_rc - sum_{k<r} L_rk U_kc, then 1 / U_rr on the diagonal or
// L_cr = (A_rc - sum_{k<r} L_ck U_kr) / U_rr below it
template <class T>
__global__ void k_ilu_factor_level(IluDev g, const int32_t* __restrict__ cols, int cnt, const T* __restrict__ av,
                                   T* __restrict__ Ux, T* __restrict__ Lx, T* __restrict__ Di) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int col = cols[t];
  for (int idx = g.Up[col]; idx < g.Up[col + 1]; ++idx) {
    const int row = g.Ui[idx];
    const T a = av[idx];
    const T u = a - ldotu(g, Lx, Ux, row, col, row);
    Ux[idx] = u;
    if (row == col) {
      Di[row] = T(1) / u;
      Lx[idx] = T(1);
    } else {
      Lx[idx] = (a - ldotu(g, Lx, Ux, col, row, row)) * Di[row];
    }
  }
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kIluThreads / 64; ++i) r += sh[i];
  return r;
}

// nonlinear residual terms (:801-824): |A - (L U)| at every U entry and at
// every strictly-lower L entry; part[block] = the block's sum
template <class T>
__global__ void __launch_bounds__(kIluThreads) k_ilu_nlres(IluDev g, const T* __restrict__ av,
                                                           const T* __restrict__ Lx, const T* __restrict__ Ux,
                                                           double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double sh[kIluThreads / 64];
  const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  double t = 0.0;
  if (idx < g.nnz) {
    const int row = g.Ui[idx], col = g.col[idx];
    const T a = av[idx];
    t = static_cast<double>(fabs(a - ldotu(g, Lx, Ux, row, col, row + 1)));
    if (row < col) t += static_cast<double>(fabs(a - ldotu(g, Lx, Ux, col, row, row + 1)));
  }
  const double s = block_sum(t, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kIluThreads) k_ilu_sum(const double* __restrict__ part, int np,
                                                         double* __restrict__ out) {
  __shared__ double sh[kIluThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = s;
}

// L y = b, rows of one level; y overwrites b (row r of L: the entries of
// column r of U above the diagonal, which is the column's last entry). The
// entries are read 8 at a time, all loads of a group in flight together,
// and summed in order.
constexpr int kIluGroup = 8;
template <class T>
__global__ void k_ilu_lower(IluDev g, const int32_t* __restrict__ rows, int cnt, const T* __restrict__ Lx,
                            double* __restrict__ x) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt || (g.gate != nullptr && *g.gate == 0)) return;
  const int r = rows[t];
  const int p0 = g.Up[r], p1 = g.Up[r + 1] - 1;
  double sum = 0.0;
  for (int base = p0; base < p1; base += kIluGroup) {
    int c[kIluGroup];
    T l[kIluGroup];
#pragma unroll
    for (int k = 0; k < kIluGroup; ++k) {
      const int idx = min(base + k, p1 - 1);
      c[k] = g.Ui[idx];
      l[k] = Lx[idx];
    }
    double xv[kIluGroup];
#pragma unroll
    for (int k = 0; k < kIluGroup; ++k) xv[k] = x[c[k]];
#pragma unroll
    for (int k = 0; k < kIluGroup; ++k)
      if (base + k < p1) sum += static_cast<double>(l[k]) * xv[k];
  }
  x[r] = x[r] - sum;
}

// U x = y, rows of one level, in place: x_r = (y_r - U_rc x_c for c > r,
// descending) / U_rr, the entries read 8 at a time
template <class T>
__global__ void k_ilu_upper(IluDev g, const int32_t* __restrict__ rows, int cnt, const T* __restrict__ Ux,
                            const T* __restrict__ Di, double* __restrict__ x) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt || (g.gate != nullptr && *g.gate == 0)) return;
  const int r = rows[t];
  const int p0 = g.ur_ptr[r], p1 = g.ur_ptr[r + 1];
  double v = x[r];
  for (int top = p1; top > p0; top -= kIluGroup) {
    int c[kIluGroup], e[kIluGroup];
#pragma unroll
    for (int k = 0; k < kIluGroup; ++k) {
      const int q = max(top - 1 - k, p0);
      c[k] = g.ur_col[q];
      e[k] = g.ur_idx[q];
    }
    double uv[kIluGroup], xv[kIluGroup];
#pragma unroll
    for (int k = 0; k < kIluGroup; ++k) {
      uv[k] = static_cast<double>(Ux[e[k]]);
      xv[k] = x[c[k]];
    }
#pragma unroll
    for (int k = 0; k < kIluGroup; ++k)
      if (top - 1 - k >= p0) v -= uv[k] * xv[k];
  }
  x[r] = v * static_cast<double>(Di[r]);
}

inline unsigned blocks(int64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

template <class T>
void ilu_init(const IluDev& g, const double* ax, T* av, T* Ux, T* Lx, T* Di, hipStream_t st) {
  if (g.nnz > 0)
    hipLaunchKernelGGL(k_ilu_init<T>, dim3(blocks(g.nnz, kIluThreads)), dim3(kIluThreads), 0, st, g, ax, av, Ux, Lx,
                       Di);
  MPIR_HIP(hipGetLastError());
}

template <class T>
void ilu_factor(const IluDev& g, const IluLevels& lv, const T* av, T* Ux, T* Lx, T* Di, hipStream_t st) {
  for (size_t l = 0; l + 1 < lv.h_lo_ptr.size(); ++l) {
    const int c0 = lv.h_lo_ptr[l], c1 = lv.h_lo_ptr[l + 1];
    hipLaunchKernelGGL(k_ilu_factor_level<T>, dim3(blocks(c1 - c0, 64)), dim3(64), 0, st, g, g.lo_rows + c0, c1 - c0,
                       av, Ux, Lx, Di);
  }
  MPIR_HIP(hipGetLastError());
}

template <class T>
void ilu_nlres(const IluDev& g, const T* av, const T* Lx, const T* Ux, double* part, double* out, hipStream_t st) {
  const unsigned nb = blocks(g.nnz, kIluThreads);
  if (nb > 0) hipLaunchKernelGGL(k_ilu_nlres<T>, dim3(nb), dim3(kIluThreads), 0, st, g, av, Lx, Ux, part);
  hipLaunchKernelGGL(k_ilu_sum, dim3(1), dim3(kIluThreads), 0, st, part, static_cast<int>(nb), out);
  MPIR_HIP(hipGetLastError());
}

template <class T>
void ilu_solve(const IluDev& g, const IluLevels& lv, const T* Lx, const T* Ux, const T* Di, double* x, hipStream_t st) {
  for (size_t l = 0; l + 1 < lv.h_lo_ptr.size(); ++l) {
    const int c0 = lv.h_lo_ptr[l], c1 = lv.h_lo_ptr[l + 1];
    hipLaunchKernelGGL(k_ilu_lower<T>, dim3(blocks(c1 - c0, 64)), dim3(64), 0, st, g, g.lo_rows + c0, c1 - c0, Lx, x);
  }
  for (size_t l = 0; l + 1 < lv.h_up_ptr.size(); ++l) {
    const int c0 = lv.h_up_ptr[l], c1 = lv.h_up_ptr[l + 1];
    hipLaunchKernelGGL(k_ilu_upper<T>, dim3(blocks(c1 - c0, 64)), dim3(64), 0, st, g, g.up_rows + c0, c1 - c0, Ux, Di,
                       x);
  }
  MPIR_HIP(hipGetLastError());
}

template void ilu_init<float>(const IluDev&, const double*, float*, float*, float*, float*, hipStream_t);
template void ilu_init<double>(const IluDev&, const double*, double*, double*, double*, double*, hipStream_t);
template void ilu_factor<float>(const IluDev&, const IluLevels&, const float*, float*, float*, float*, hipStream_t);
template void ilu_factor<double>(const IluDev&, const IluLevels&, const double*, double*, double*, double*,
                                 hipStream_t);
template void ilu_nlres<float>(const IluDev&, const float*, const float*, const float*, double*, double*, hipStream_t);
template void ilu_nlres<double>(const IluDev&, const double*, const double*, const double*, double*, double*,
                                hipStream_t);
template void ilu_solve<float>(const IluDev&, const IluLevels&, const float*, const float*, const float*, double*,
                               hipStream_t);
template void ilu_solve<double>(const IluDev&, const IluLevels&, const double*, const double*, const double*, double*,
                                hipStream_t);

}  // namespace mpir
