// Multifrontal LDL^T for quasi-definite matrices on gfx950.
//
// Numerically this is QDLDL_factor (include/MPIR/qdldl.hpp:172-345 of
// yex33/MPIR): no pivoting, L unit lower triangular, D diagonal, each L entry
// formed as y * Dinv with Dinv = 1 / D (qdldl.hpp:315, :339), and a zero
// pivot reported as failure (qdldl.hpp:331-333). The loop order is different:
// columns are grouped into supernodes and eliminated front by front, level by
// level of the supernodal tree, which is what exposes parallelism.
//
// Storage (per supernode s with nc columns and nr rows below the block,
// m = nc + nr, ld = panel_ld(nc, nr)):
//   panel  ld x nc column-major: rows [cols of s, R_s] and, for big fronts
//          (m > 64), nc more rows that start as the identity. On exit the
//          strictly-lower part of the first m rows holds L, the diagonal D,
//          and the extra rows Z = L_11^{-T} D^{-1} (upper triangular): the
//          blocked elimination applied to [F; I] computes them with the same
//          panel solves and updates, since eliminating the first nc columns
//          of [[F_11, F_21^T, I], [F_21, F_22, 0], [I, 0, 0]] leaves
//          I * L_11^{-T} D^{-1} in the rows of the identity
//   U_s    nr x nr column-major (ld = nr), lower part: the update matrix
//          handed to the parent (extend-add through relidx)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common.hpp"
#include "../plan.hpp"
#include "kernels.hpp"

namespace mpir {

namespace {

__device__ __forceinline__ void flag_zero_pivot(int* err) { atomicOr(err, 1); }

// ---------------------------------------------------------------- scatter
template <class T>
__global__ void k_scatter_values(int64_t nnz, const double* __restrict__ ax,
                                 const int64_t* __restrict__ dst, T* __restrict__ panel) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < nnz) panel[dst[q]] = static_cast<T>(ax[q]);
}

// ---------------------------------------------------------------- small fronts
// One wavefront per front (m <= 64), the whole front in LDS.
template <class T>
__global__ void __launch_bounds__(64) k_small_fronts(FactorArgs<T> a, const int32_t* __restrict__ list) {
  __shared__ T F[kSmallFront][kSmallFront + 1];
  __shared__ T lb[kSmallFront];
  const int s = list[blockIdx.x];
  const int lane = threadIdx.x;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  T* P = a.panel + a.panel_off[s];
  T* U = nr > 0 ? a.upool + a.upd_off[s] : nullptr;

  // load the panel columns, zero the update block
  for (int c = 0; c < m; ++c) {
    if (lane < m) F[lane][c] = (c < nc) ? P[lane + static_cast<int64_t>(c) * m] : T(0);
  }
  __syncthreads();
  // extend-add the children's update matrices, one child at a time
  for (int t = a.child_ptr[s]; t < a.child_ptr[s + 1]; ++t) {
    const int c = a.child_list[t];
    const int nrc = a.nrow[c];
    if (nrc == 0) continue;
    const T* Uc = a.upool + a.upd_off[c];
    const int32_t* rel = a.relidx + a.rows_off[c];
    const int ra = lane < nrc ? rel[lane] : 0;
    for (int b = 0; b < nrc; ++b) {
      const int rb = rel[b];
      if (lane >= b && lane < nrc) F[ra][rb] += Uc[lane + static_cast<int64_t>(b) * nrc];
    }
    __syncthreads();
  }
  // right-looking elimination of the nc panel columns
  for (int j = 0; j < nc; ++j) {
    const T d = F[j][j];
    if (d == T(0)) flag_zero_pivot(a.err);
    const T dinv = T(1) / d;
    if (lane > j && lane < m) lb[lane] = F[lane][j] * dinv;
    __syncthreads();
    if (lane > j && lane < m) {
      const T y = F[lane][j];
      for (int c = j + 1; c <= lane; ++c) F[lane][c] -= y * lb[c];
      F[lane][j] = lb[lane];
    }
    if (lane == 0) {
      a.D[f + j] = d;
      a.Dinv[f + j] = dinv;
    }
    __syncthreads();
  }
  for (int c = 0; c < nc; ++c)
    if (lane < m && lane >= c) P[lane + static_cast<int64_t>(c) * m] = F[lane][c];
  for (int b = 0; b < nr; ++b)
    if (lane >= b && lane < nr) U[lane + static_cast<int64_t>(b) * nr] = F[nc + lane][nc + b];
}

// ---------------------------------------------------------------- assembly
// Block ownership: the task (s, c0, c1, t0) owns the entries of the front
// in columns [c0, c1) and rows [t0, t0 + kAsmRows): it zeroes them in the
// update matrix and adds every child's contribution that lands there,
// children in fixed order, so no two tasks write the same entry. Child
// columns (rows) landing in a range of the parent are contiguous because the
// relative indices increase.
template <class T>
__global__ void __launch_bounds__(256) k_assemble(FactorArgs<T> a, const Task4* __restrict__ tasks,
                                                  const int32_t* __restrict__ cptr, const int32_t* __restrict__ cl) {
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, c0 = tk.b, c1 = tk.c, t0 = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int t1 = min(m, t0 + kAsmRows);
  T* P = a.panel + a.panel_off[s];
  T* U = nr > 0 ? a.upool + a.upd_off[s] : nullptr;
  const int tid = threadIdx.x;
  for (int c = c0 < nc ? nc : c0; c < c1; ++c) {
    const int cb = c - nc;
    for (int r = max(c, t0) + tid; r < t1; r += blockDim.x) U[(r - nc) + static_cast<int64_t>(cb) * nr] = T(0);
  }
  __syncthreads();
  for (int e = cptr[blockIdx.x]; e < cptr[blockIdx.x + 1]; ++e) {
    const int ch = cl[5 * e], b0 = cl[5 * e + 1], b1 = cl[5 * e + 2], a0 = cl[5 * e + 3], a1 = cl[5 * e + 4];
    const int nrc = a.nrow[ch];
    const T* Uc = a.upool + a.upd_off[ch];
    const int32_t* rel = a.relidx + a.rows_off[ch];
    for (int b = b0; b < b1; ++b) {
      const int rb = rel[b];
      const T* src = Uc + static_cast<int64_t>(b) * nrc;
      for (int q = max(b, a0) + tid; q < a1; q += blockDim.x) {
        const int ra = rel[q];
        if (rb < nc) P[ra + rb * ld] += src[q];
        else U[(ra - nc) + static_cast<int64_t>(rb - nc) * nr] += src[q];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- big fronts
// Diagonal block of step k: LDL^T of an nbk x nbk block by one wavefront,
// lane r holding row r of the block in registers; column j's multipliers are
// broadcast through LDS.
template <class T>
__global__ void __launch_bounds__(64) k_diag_block(FactorArgs<T> a, const Task4* __restrict__ fronts) {
  __shared__ T lb[kPanelNB];
  const Task4 tk = fronts[blockIdx.x];
  const int s = tk.a, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  const int64_t ld = panel_ld(nc, a.nrow[s]);
  T* P = a.panel + a.panel_off[s] + k0 + k0 * ld;
  const int r = threadIdx.x;
  T b[kPanelNB];
#pragma unroll
  for (int c = 0; c < kPanelNB; ++c) b[c] = (c < nbk && r < nbk) ? P[r + c * ld] : T(0);
#pragma unroll
  for (int j = 0; j < kPanelNB; ++j) {
    if (j < nbk) {
      const T d = __shfl(b[j], j);
      const T dinv = T(1) / d;
      const T y = b[j];
      const T l = y * dinv;
      lb[r] = l;
      __syncthreads();
      if (r > j) {
#pragma unroll
        for (int c = j + 1; c < kPanelNB; ++c)
          if (c <= r) b[c] -= y * lb[c];
        b[j] = l;
      }
      if (r == j) {
        if (d == T(0)) flag_zero_pivot(a.err);
        a.D[f + k0 + j] = d;
        a.Dinv[f + k0 + j] = dinv;
      }
      __syncthreads();
    }
  }
  if (r < nbk) {
#pragma unroll
    for (int c = 0; c < kPanelNB; ++c)
      if (c <= r && c < nbk) P[r + c * ld] = b[c];
  }
}

// Rows below the diagonal block: X L_kk^T = A_rows, L_rows = X Dinv.
// One thread per row, the row in registers, eliminated right-looking so the
// updates of later columns are independent of each other. Rows r >= m are
// the identity rows: row m + i enters at the step of its own column i with
// e_i as its value (it is zero in every earlier column).
template <class T>
__global__ void __launch_bounds__(256) k_panel_trsm(FactorArgs<T> a, const Task4* __restrict__ tasks) {
  __shared__ T Lk[kPanelNB][kPanelNB + 1];
  __shared__ T dv[kPanelNB];
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, r0 = tk.b, r1 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int64_t ld = panel_ld(nc, a.nrow[s]);
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  T* P = a.panel + a.panel_off[s];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  for (int c = w; c < nbk; c += 4)
    if (lane < nbk && lane > c) Lk[lane][c] = P[(k0 + lane) + (k0 + c) * ld];
  if (tid < nbk) dv[tid] = a.Dinv[f + k0 + tid];
  __syncthreads();
  const int r = r0 + tid;
  if (r >= r1) return;
  const int unit = r >= m + k0 ? r - m - k0 : -1;  // identity row of this block
  T x[kPanelNB];
#pragma unroll
  for (int c = 0; c < kPanelNB; ++c)
    x[c] = c >= nbk ? T(0) : unit >= 0 ? (c == unit ? T(1) : T(0)) : P[r + (k0 + c) * ld];
#pragma unroll
  for (int t = 0; t < kPanelNB; ++t) {
    if (t < nbk) {
      const T xt = x[t];
#pragma unroll
      for (int c = t + 1; c < kPanelNB; ++c) x[c] -= xt * Lk[c][t];
    }
  }
#pragma unroll
  for (int c = 0; c < kPanelNB; ++c)
    if (c < nbk) P[r + (k0 + c) * ld] = x[c] * dv[c];
}

// Trailing update tile: C[r0:r0+64, c0:c0+64] -= L_r D L_c^T over the nbk
// columns of step k, lower part only. Columns < nc live in the panel,
// columns >= nc in the update matrix. Tiles with r0 >= m cover identity rows
// (rows m .. m+nc) and only panel columns.
template <class T>
__global__ void __launch_bounds__(256) k_update_tiles(FactorArgs<T> a, const Task4* __restrict__ tasks) {
  __shared__ T As[kPanelNB][kTile + 1];
  __shared__ T Bs[kPanelNB][kTile + 1];
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, r0 = tk.b, c0 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int rlim = r0 >= m ? m + nc : m;
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  const T* P = a.panel + a.panel_off[s];
  T* Pw = a.panel + a.panel_off[s];
  T* U = nr > 0 ? a.upool + a.upd_off[s] : nullptr;
  const int tid = threadIdx.x;
  for (int e = tid; e < nbk * kTile; e += blockDim.x) {
    const int i = e % kTile, t = e / kTile;
    const int ra = r0 + i, rb = c0 + i;
    As[t][i] = ra < rlim ? P[ra + (k0 + t) * ld] : T(0);
    Bs[t][i] = rb < m ? P[rb + (k0 + t) * ld] * a.D[f + k0 + t] : T(0);
  }
  __syncthreads();
  const int ti = (tid % 16) * 4, tj = (tid / 16) * 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
  for (int t = 0; t < nbk; ++t) {
    T av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) { av[i] = As[t][ti + i]; bv[i] = Bs[t][tj + i]; }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = c0 + tj + j;
    if (c >= m) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + ti + i;
      if (r >= rlim || r < c) continue;
      if (c < nc) Pw[r + c * ld] -= acc[i][j];
      else if (r < m) U[(r - nc) + static_cast<int64_t>(c - nc) * nr] -= acc[i][j];
    }
  }
}

template <class T>
__global__ void k_count_positive(int n, const T* __restrict__ D, int* __restrict__ out) {
  int cnt = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    cnt += D[i] > T(0) ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, cnt);
}

inline unsigned blocks_for(int64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

template <class T>
void factor_numeric(const FactorArgs<T>& a, const DeviceFactorPlan& dp, const double* ax_scaled,
                    int64_t nnz, hipStream_t st) {
  MPIR_HIP(hipMemsetAsync(a.panel, 0, sizeof(T) * dp.panel_entries, st));
  MPIR_HIP(hipMemsetAsync(a.err, 0, sizeof(int), st));
  if (nnz > 0)
    hipLaunchKernelGGL(k_scatter_values<T>, dim3(blocks_for(nnz, 256)), dim3(256), 0, st, nnz, ax_scaled,
                       dp.val_dst, a.panel);
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL(k_small_fronts<T>, dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0);
    const int a0 = dp.h_asm_ptr[l], a1 = dp.h_asm_ptr[l + 1];
    if (a1 > a0)
      hipLaunchKernelGGL(k_assemble<T>, dim3(a1 - a0), dim3(256), 0, st, a, dp.asm_task + a0, dp.asm_child_ptr + a0,
                         dp.asm_child);
    const int st0 = dp.h_level_step_ptr[l], st1 = dp.h_level_step_ptr[l + 1];
    for (int q = st0; q < st1; ++q) {
      const int d0 = dp.h_diag_ptr[q], d1 = dp.h_diag_ptr[q + 1];
      hipLaunchKernelGGL(k_diag_block<T>, dim3(d1 - d0), dim3(64), 0, st, a, dp.diag_task + d0);
      const int p0 = dp.h_pan_ptr[q], p1 = dp.h_pan_ptr[q + 1];
      if (p1 > p0) hipLaunchKernelGGL(k_panel_trsm<T>, dim3(p1 - p0), dim3(256), 0, st, a, dp.pan_task + p0);
      const int u0 = dp.h_upd_ptr[q], u1 = dp.h_upd_ptr[q + 1];
      if (u1 > u0) hipLaunchKernelGGL(k_update_tiles<T>, dim3(u1 - u0), dim3(256), 0, st, a, dp.upd_task + u0);
    }
  }
  MPIR_HIP(hipGetLastError());
}

template <class T>
void invert_diag_blocks(const FactorArgs<T>& a, const DeviceFactorPlan& dp, hipStream_t st) {
  for (int l = 0; l < dp.nlevels; ++l) {
    const int st0 = dp.h_level_step_ptr[l], st1 = dp.h_level_step_ptr[l + 1];
    for (int q = st0; q < st1; ++q) {
      const int p0 = dp.h_pan_aug_ptr[q], p1 = dp.h_pan_ptr[q + 1];
      if (p1 > p0) hipLaunchKernelGGL(k_panel_trsm<T>, dim3(p1 - p0), dim3(256), 0, st, a, dp.pan_task + p0);
      const int u0 = dp.h_upd_aug_ptr[q], u1 = dp.h_upd_ptr[q + 1];
      if (u1 > u0) hipLaunchKernelGGL(k_update_tiles<T>, dim3(u1 - u0), dim3(256), 0, st, a, dp.upd_task + u0);
    }
  }
  MPIR_HIP(hipGetLastError());
}

template <class T>
__global__ void k_reciprocal(int n, const T* __restrict__ x, T* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = T(1) / x[i];
}

template <class T>
void reciprocal(int n, const T* x, T* y, hipStream_t st) {
  hipLaunchKernelGGL(k_reciprocal<T>, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, x, y);
}

template <class T>
void count_positive(int n, const T* D, int* out, hipStream_t st) {
  MPIR_HIP(hipMemsetAsync(out, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_count_positive<T>, dim3(std::min<unsigned>(1024, blocks_for(n, 256))), dim3(256), 0, st, n, D,
                     out);
}

template void factor_numeric<double>(const FactorArgs<double>&, const DeviceFactorPlan&, const double*, int64_t,
                                     hipStream_t);
template void factor_numeric<float>(const FactorArgs<float>&, const DeviceFactorPlan&, const double*, int64_t,
                                    hipStream_t);
template void invert_diag_blocks<double>(const FactorArgs<double>&, const DeviceFactorPlan&, hipStream_t);
template void invert_diag_blocks<float>(const FactorArgs<float>&, const DeviceFactorPlan&, hipStream_t);
template void reciprocal<double>(int, const double*, double*, hipStream_t);
template void reciprocal<float>(int, const float*, float*, hipStream_t);
template void count_positive<double>(int, const double*, int*, hipStream_t);
template void count_positive<float>(int, const float*, int*, hipStream_t);

}  // namespace mpir
