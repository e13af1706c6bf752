// Supernodal triangular solves with the multifrontal factor:
//   forward  (L + I) y = b,  y := D^{-1} y,  backward (L + I)^T x = y.
// This is QDLDL_solve (include/MPIR/qdldl.hpp:417-432 of yex33/MPIR: Lsolve
// :362-372, the Dinv step :430, Ltsolve :389-400) with the factor held in
// factor precision TF and every product and sum done in the precision of x,
// TX, as QDLDL_Lsolve/QDLDL_Ltsolve do with XPREC(Lx[j]).
//
// Forward: a front gathers b on its columns, adds its children's update
// vectors (extend-add), solves its unit-lower diagonal block and leaves
// u_s = v_R - L_21 y on its rows below for the parent. Backward: a front
// gathers x on R_s (all ancestors, already final) and solves its block
// transposed. Fronts of one level are independent.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common.hpp"
#include "kernels.hpp"

namespace mpir {

namespace {

// ---------------------------------------------------------------- small fronts
template <class TF, class TX>
__global__ void __launch_bounds__(64) k_fwd_small(SolveArgs<TF> a, const int32_t* __restrict__ list, TX* __restrict__ x,
                                                  TX* __restrict__ uvec) {
  __shared__ TX v[kSmallFront];
  const int s = list[blockIdx.x];
  const int lane = threadIdx.x;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const TF* P = a.panel + a.panel_off[s];
  if (lane < m) v[lane] = lane < nc ? x[f + lane] : TX(0);
  __syncthreads();
  for (int t = a.child_ptr[s]; t < a.child_ptr[s + 1]; ++t) {
    const int c = a.child_list[t];
    const int nrc = a.nrow[c];
    if (lane < nrc) v[a.relidx[a.rows_off[c] + lane]] += uvec[a.rows_off[c] + lane];
    __syncthreads();
  }
  for (int j = 0; j < nc; ++j) {
    const TX yj = v[j];
    if (lane > j && lane < m) v[lane] -= static_cast<TX>(P[lane + static_cast<int64_t>(j) * m]) * yj;
    __syncthreads();
  }
  if (lane < nc) x[f + lane] = v[lane] * static_cast<TX>(a.Dinv[f + lane]);
  else if (lane < m) uvec[a.rows_off[s] + (lane - nc)] = v[lane];
}

template <class TF, class TX>
__global__ void __launch_bounds__(64) k_bwd_small(SolveArgs<TF> a, const int32_t* __restrict__ list, TX* __restrict__ x) {
  __shared__ TX w[kSmallFront];
  const int s = list[blockIdx.x];
  const int lane = threadIdx.x;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const TF* P = a.panel + a.panel_off[s];
  if (lane < nc) w[lane] = x[f + lane];
  else if (lane < m) w[lane] = x[a.rows[a.rows_off[s] + (lane - nc)]];
  __syncthreads();
  // lane j < nc: z_j = w_j - sum over rows r >= nc of L[r, j] w_r
  if (lane < nc) {
    TX z = w[lane];
    for (int r = nc; r < m; ++r) z -= static_cast<TX>(P[r + static_cast<int64_t>(lane) * m]) * w[r];
    w[lane] = z;
  }
  __syncthreads();
  // unit upper solve with L_11^T, bottom to top
  for (int j = nc - 1; j >= 0; --j) {
    const TX xj = w[j];
    if (lane < j) w[lane] -= static_cast<TX>(P[j + static_cast<int64_t>(lane) * m]) * xj;
    __syncthreads();
  }
  if (lane < nc) x[f + lane] = w[lane];
}

// ---------------------------------------------------------------- big fronts
// Front vector v = [x(cols of s); u_s]: zero u_s, add children's updates.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_fwd_big_init(SolveArgs<TF> a, const int32_t* __restrict__ list,
                                                      TX* __restrict__ x, TX* __restrict__ uvec) {
  const int s = list[blockIdx.x];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  TX* us = uvec + a.rows_off[s];
  for (int r = threadIdx.x; r < nr; r += blockDim.x) us[r] = TX(0);
  __syncthreads();
  for (int t = a.child_ptr[s]; t < a.child_ptr[s + 1]; ++t) {
    const int c = a.child_list[t];
    const int nrc = a.nrow[c];
    const int32_t* rel = a.relidx + a.rows_off[c];
    const TX* uc = uvec + a.rows_off[c];
    for (int r = threadIdx.x; r < nrc; r += blockDim.x) {
      const int pos = rel[r];
      if (pos < nc) x[f + pos] += uc[r];
      else us[pos - nc] += uc[r];
    }
    __syncthreads();
  }
}

// Diagonal block of step k (forward): unit lower solve in LDS.
template <class TF, class TX>
__global__ void __launch_bounds__(64) k_fwd_diag(SolveArgs<TF> a, const Task4* __restrict__ fronts, TX* __restrict__ x) {
  __shared__ TX v[kPanelNB];
  const Task4 tk = fronts[blockIdx.x];
  const int s = tk.a, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  const TF* P = a.panel + a.panel_off[s];
  const int lane = threadIdx.x;
  if (lane < nbk) v[lane] = x[f + k0 + lane];
  __syncthreads();
  for (int j = 0; j < nbk; ++j) {
    const TX yj = v[j];
    if (lane > j && lane < nbk) v[lane] -= static_cast<TX>(P[(k0 + lane) + static_cast<int64_t>(k0 + j) * m]) * yj;
    __syncthreads();
  }
  if (lane < nbk) x[f + k0 + lane] = v[lane];
}

// Rows below the diagonal block of step k: v_r -= L[r, block] y_block.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_fwd_gemv(SolveArgs<TF> a, const Task4* __restrict__ tasks, TX* __restrict__ x,
                                                  TX* __restrict__ uvec) {
  __shared__ TX y[kPanelNB];
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, r0 = tk.b, r1 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  const TF* P = a.panel + a.panel_off[s];
  if (threadIdx.x < nbk) y[threadIdx.x] = x[f + k0 + threadIdx.x];
  __syncthreads();
  const int r = r0 + threadIdx.x;
  if (r >= r1) return;
  TX acc = TX(0);
  for (int t = 0; t < nbk; ++t) acc += static_cast<TX>(P[r + static_cast<int64_t>(k0 + t) * m]) * y[t];
  TX* dst = r < nc ? x + f + r : uvec + a.rows_off[s] + (r - nc);
  *dst -= acc;
}

template <class TF, class TX>
__global__ void __launch_bounds__(256) k_fwd_big_fin(SolveArgs<TF> a, const int32_t* __restrict__ list, TX* __restrict__ x) {
  const int s = list[blockIdx.x];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  for (int r = threadIdx.x; r < nc; r += blockDim.x) x[f + r] *= static_cast<TX>(a.Dinv[f + r]);
}

// Backward, rows below the block of step k: partial column sums
// part[task][c] = sum over the task's rows r of L[r, k0 + c] w_r.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_bwd_gemv(SolveArgs<TF> a, const Task4* __restrict__ tasks, const TX* __restrict__ x,
                                                  TX* __restrict__ part) {
  __shared__ TX red[4][kPanelNB];
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, r0 = tk.b, r1 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  const TF* P = a.panel + a.panel_off[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = r0 + threadIdx.x;
  TX wr = TX(0);
  if (r < r1) wr = r < nc ? x[f + r] : x[a.rows[a.rows_off[s] + (r - nc)]];
  for (int c = 0; c < nbk; ++c) {
    TX pv = r < r1 ? static_cast<TX>(P[r + static_cast<int64_t>(k0 + c) * m]) * wr : TX(0);
    for (int off = 32; off > 0; off >>= 1) pv += __shfl_down(pv, off);
    if (lane == 0) red[wv][c] = pv;
  }
  __syncthreads();
  if (threadIdx.x < nbk) {
    const int c = threadIdx.x;
    part[static_cast<int64_t>(blockIdx.x) * kPanelNB + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
  }
}

// Backward diagonal block: z = x_block - sum of partials, then unit upper
// solve with L_kk^T.
template <class TF, class TX>
__global__ void __launch_bounds__(64) k_bwd_diag(SolveArgs<TF> a, const Task4* __restrict__ fronts, int task_base,
                                                 TX* __restrict__ x, const TX* __restrict__ part) {
  __shared__ TX w[kPanelNB];
  const Task4 tk = fronts[blockIdx.x];
  const int s = tk.a, t0 = tk.b, t1 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  const TF* P = a.panel + a.panel_off[s];
  const int lane = threadIdx.x;
  if (lane < nbk) {
    TX z = x[f + k0 + lane];
    for (int t = t0; t < t1; ++t) z -= part[static_cast<int64_t>(t - task_base) * kPanelNB + lane];
    w[lane] = z;
  }
  __syncthreads();
  for (int j = nbk - 1; j >= 0; --j) {
    const TX xj = w[j];
    if (lane < j) w[lane] -= static_cast<TX>(P[(k0 + j) + static_cast<int64_t>(k0 + lane) * m]) * xj;
    __syncthreads();
  }
  if (lane < nbk) x[f + k0 + lane] = w[lane];
}

}  // namespace

template <class TF, class TX>
void ldl_solve(const SolveArgs<TF>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st) {
  TX* part = reinterpret_cast<TX*>(dp.scratch);
  // forward, leaves to root
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_fwd_small<TF, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x, uvec);
    const int b0 = dp.h_big_ptr[l], b1 = dp.h_big_ptr[l + 1];
    if (b1 == b0) continue;
    hipLaunchKernelGGL((k_fwd_big_init<TF, TX>), dim3(b1 - b0), dim3(256), 0, st, a, dp.big_list + b0, x, uvec);
    for (int q = dp.h_level_step_ptr[l]; q < dp.h_level_step_ptr[l + 1]; ++q) {
      const int d0 = dp.h_diag_ptr[q], d1 = dp.h_diag_ptr[q + 1];
      hipLaunchKernelGGL((k_fwd_diag<TF, TX>), dim3(d1 - d0), dim3(64), 0, st, a, dp.diag_task + d0, x);
      const int p0 = dp.h_task_ptr[q], p1 = dp.h_task_ptr[q + 1];
      if (p1 > p0)
        hipLaunchKernelGGL((k_fwd_gemv<TF, TX>), dim3(p1 - p0), dim3(256), 0, st, a, dp.gemv_task + p0, x, uvec);
    }
    hipLaunchKernelGGL((k_fwd_big_fin<TF, TX>), dim3(b1 - b0), dim3(256), 0, st, a, dp.big_list + b0, x);
  }
  // backward, root to leaves
  for (int l = dp.nlevels - 1; l >= 0; --l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0) hipLaunchKernelGGL((k_bwd_small<TF, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x);
    for (int q = dp.h_level_step_ptr[l + 1] - 1; q >= dp.h_level_step_ptr[l]; --q) {
      const int p0 = dp.h_task_ptr[q], p1 = dp.h_task_ptr[q + 1];
      if (p1 > p0)
        hipLaunchKernelGGL((k_bwd_gemv<TF, TX>), dim3(p1 - p0), dim3(256), 0, st, a, dp.gemv_task + p0, x, part);
      const int d0 = dp.h_diag_ptr[q], d1 = dp.h_diag_ptr[q + 1];
      hipLaunchKernelGGL((k_bwd_diag<TF, TX>), dim3(d1 - d0), dim3(64), 0, st, a, dp.diag_task + d0, p0, x, part);
    }
  }
  MPIR_HIP(hipGetLastError());
}

template void ldl_solve<double, double>(const SolveArgs<double>&, const DeviceSolvePlan&, double*, double*, hipStream_t);
template void ldl_solve<float, double>(const SolveArgs<float>&, const DeviceSolvePlan&, double*, double*, hipStream_t);
template void ldl_solve<float, float>(const SolveArgs<float>&, const DeviceSolvePlan&, float*, float*, hipStream_t);

}  // namespace mpir
