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
//          strictly-lower part of the first m rows holds L (small fronts
//          also keep D on the diagonal; big fronts keep the strictly-lower
//          part of each 64 x 64 diagonal block transposed, in the block's
//          upper triangle), and the extra rows Z = L_11^{-T} D^{-1} (upper
//          triangular): the
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
// Panel step k of big front s, task (s, r0, r1, k). Every workgroup factors
// the nbk x nbk diagonal block in LDS (right-looking LDL^T) and carries 64
// identity rows through the same elimination, which leaves
// Z = L_kk^{-T} D_k^{-1} (upper triangular) in them. The task with
// r0 = k0 owns the diagonal block and writes L_kk, D and D^{-1}; every other
// task forms its rows of L as X Z, X = its rows of the current panel
// (identity rows of the front, r >= m, enter at their own column as e_i).
// refactor = false: L_kk and D are already final (a factor loaded from
// outside); the block is rebuilt from them and only Z is formed.
template <class T>
__global__ void __launch_bounds__(256) k_panel_step(FactorArgs<T> a, const Task4* __restrict__ tasks, bool refactor) {
  __shared__ T F[kPanelNB][kPanelNB + 1];  // diagonal block, then the task's rows
  __shared__ T G[kPanelNB][kPanelNB + 1];  // identity rows -> Z
  __shared__ T lc[kPanelNB], lg[kPanelNB];
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, r0 = tk.b, r1 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int64_t ld = panel_ld(nc, a.nrow[s]);
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  T* P = a.panel + a.panel_off[s];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool owner = r0 == k0;
  const int rows = r1 - r0;
  // this thread's part of its row (columns 16w .. 16w + 15), loaded early
  const int r = r0 + lane;
  const int unit = r >= m + k0 ? r - m - k0 : -1;
  T xv[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int c = 16 * w + u;
    xv[u] = (owner || lane >= rows || c >= nbk) ? T(0)
            : unit >= 0                         ? (c == unit ? T(1) : T(0))
                                                : P[r + (k0 + c) * ld];
  }
  const T* Dk = a.D + f + k0;
  for (int c = w; c < kPanelNB; c += 4) {
    T v = T(0);
    if (lane < nbk && c < nbk && lane >= c) {
      if (refactor) v = P[(k0 + lane) + (k0 + c) * ld];
      else v = lane == c ? Dk[c] : P[(k0 + c) + (k0 + lane) * ld] * Dk[c];
    }
    F[lane][c] = v;
    G[lane][c] = lane == c ? T(1) : T(0);
  }
  __syncthreads();
  for (int j = 0; j < nbk; ++j) {
    const T d = F[j][j];
    const T dinv = T(1) / d;
    if (w == 0) {
      if (lane > j && lane < nbk) {
        const T l = F[lane][j] * dinv;
        lc[lane] = l;
        // L_kk goes to the block's upper triangle (transposed): the other
        // workgroups of this launch are still reading the lower triangle
        if (owner && refactor) P[(k0 + j) + (k0 + lane) * ld] = l;
      }
      if (lane == j && owner && refactor) {
        if (d == T(0)) flag_zero_pivot(a.err);
        a.D[f + k0 + j] = d;
        a.Dinv[f + k0 + j] = dinv;
      }
    } else if (w == 1) {
      if (lane <= j) lg[lane] = G[lane][j] * dinv;
    }
    __syncthreads();
    // trailing block rows (lower part) and active identity rows i <= j
    if (refactor && lane > j && lane < nbk) {
      const T l = lc[lane];
      for (int c = j + 1 + w; c <= lane; c += 4) F[lane][c] -= l * F[c][j];
    }
    if (lane <= j) {
      const T g = lg[lane];
      for (int c = j + 1 + w; c < nbk; c += 4) G[lane][c] -= g * F[c][j];
      if (w == 0) G[lane][j] = g;
    }
    __syncthreads();
  }
  if (owner) return;
  // rows of L: X Z, thread (row lane, columns 16w ..)
#pragma unroll
  for (int u = 0; u < 16; ++u) F[lane][16 * w + u] = xv[u];
  __syncthreads();
  if (lane >= rows) return;
  T acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = T(0);
  for (int t = 0; t < nbk; ++t) {
    const T xt = F[lane][t];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] += xt * G[t][16 * w + u];
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int c = 16 * w + u;
    if (c < nbk) P[r + (k0 + c) * ld] = acc[u];
  }
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
      const int p0 = dp.h_pan_ptr[q], p1 = dp.h_pan_ptr[q + 1];
      if (p1 > p0) hipLaunchKernelGGL(k_panel_step<T>, dim3(p1 - p0), dim3(256), 0, st, a, dp.pan_task + p0, true);
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
      if (p1 > p0) hipLaunchKernelGGL(k_panel_step<T>, dim3(p1 - p0), dim3(256), 0, st, a, dp.pan_task + p0, false);
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
