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
// One launch per level handles every big front of the level. A front is cut
// into 64-row blocks; its diagonal blocks form a chain (block i needs y of
// every column block j < i), carried by per-block flags that hold the epoch
// of the current pass, so flags never need resetting. Workgroups are
// launched in an order where every wait targets a workgroup with a smaller
// index, which the in-order dispatch has already started. Waits are bounded:
// on expiry the kernel records an error and finishes.
constexpr int kRB = kSolveRB;
constexpr int kSpinLimit = 1 << 24;

// Values passed between workgroups of one launch use relaxed device-scope
// atomic loads and stores: on gfx950 these are plain accesses with the sc1
// (device-coherent) bit, so no cache writeback or invalidation is needed. A
// writer waits for its stores to complete (s_waitcnt) before a barrier and
// the flag store; a reader spins on the flag, then loads the data.
template <class T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wait_flag(const int* flag, int epoch, int* err) {
  int it = 0;
  while (ld_dev(flag) != epoch) {
    __builtin_amdgcn_s_sleep(1);
    if (++it > kSpinLimit) {
      atomicOr(err, 2);
      break;
    }
  }
  // keeps the compiler from moving the data loads above the wait (no cache
  // maintenance is emitted at workgroup scope)
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// every thread of the block calls this after its last data store
__device__ __forceinline__ void publish_flag(int* flag, int epoch) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) st_dev(flag, epoch);
}

// Forward, row block i of front s: v = [b on columns; 0 on R] + children's
// update vectors - sum_j L[block i, block j] y_j; diagonal blocks then solve
// their unit-lower block, write y (ybuf) and D^{-1} y (x) and publish.
// Everything that does not depend on another block (the diagonal block, the
// tile of the last dependency) is loaded before the waits.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_fwd_front(SolveArgs<TF> a, const int32_t* __restrict__ tasks,
                                                   const int32_t* __restrict__ cptr, const int32_t* __restrict__ cl,
                                                   const int32_t* __restrict__ flag_off, int* flags, int epoch,
                                                   TX* x, TX* __restrict__ uvec, TX* ybuf, int* err) {
  __shared__ TX vsh[kRB];
  __shared__ TX part[4][kRB];
  __shared__ TF Ld[kRB][kRB + 1];
  const int s = tasks[2 * blockIdx.x], i = tasks[2 * blockIdx.x + 1];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int r0 = i * kRB;
  const int rows = min(kRB, m - r0);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const TF* P = a.panel + a.panel_off[s];
  TX* us = uvec + a.rows_off[s];
  const bool diag = r0 < nc;
  const int bwd = diag ? min(kRB, nc - r0) : 0;
  // diagonal block into LDS
  for (int t = w; t < bwd; t += 4)
    if (lane < rows) Ld[lane][t] = P[(r0 + lane) + static_cast<int64_t>(r0 + t) * m];
  if (tid < kRB) {
    const int r = r0 + tid;
    vsh[tid] = (tid < rows && r < nc) ? x[f + r] : TX(0);
  }
  __syncthreads();
  for (int e = cptr[blockIdx.x]; e < cptr[blockIdx.x + 1]; ++e) {
    const int c = cl[3 * e], q0 = cl[3 * e + 1], q1 = cl[3 * e + 2];
    const int32_t* rel = a.relidx + a.rows_off[c];
    const TX* uc = uvec + a.rows_off[c];
    for (int q = q0 + tid; q < q1; q += blockDim.x) vsh[rel[q] - r0] += uc[q];
    __syncthreads();
  }
  const int ncb = (nc + kRB - 1) / kRB;
  const int jend = min(i, ncb);
  const int rr = r0 + (lane < rows ? lane : 0);
  TX acc = TX(0);
  // column blocks oldest first: the newest dependency (j = i - 1) is the last
  // item of its wave, so after its flag only one 64-column update remains;
  // each tile is loaded before the wait
  for (int j = w; j < jend; j += 4) {
    const int c0 = j * kRB;
    const int bw = min(kRB, nc - c0);
    const TF* col = P + rr + static_cast<int64_t>(c0) * m;
    TF lv[kRB];
#pragma unroll
    for (int t = 0; t < kRB; ++t) lv[t] = t < bw ? col[static_cast<int64_t>(t) * m] : TF(0);
    if (lane == 0) wait_flag(flags + flag_off[s] + j, epoch, err);
    const TX ylane = lane < bw ? ld_dev(ybuf + f + c0 + lane) : TX(0);
#pragma unroll
    for (int t = 0; t < kRB; ++t) acc += static_cast<TX>(lv[t]) * __shfl(ylane, t);
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    TX va = lane < rows ? vsh[lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) : TX(0);
    if (diag) {
      for (int t = 0; t < bwd; ++t) {
        const TX yt = __shfl(va, t);
        if (lane > t && lane < rows) va -= static_cast<TX>(Ld[lane][t]) * yt;
      }
      if (lane < bwd) {
        st_dev(ybuf + f + r0 + lane, va);
        x[f + r0 + lane] = va * static_cast<TX>(a.Dinv[f + r0 + lane]);
      } else if (lane < rows) {
        us[r0 + lane - nc] = va;
      }
    } else if (lane < rows) {
      us[r0 + lane - nc] = va;
    }
  }
  if (diag) publish_flag(flags + flag_off[s] + i, epoch);
}

// Backward, rows below the columns of a front (R rows, all final): partial
// sums part[(j * nch + c) * 64 + t] = sum over the chunk's rows r of
// L[nc + r, 64 j + t] x[R_r]. Columns are contiguous in the panel, so each
// wave walks whole columns (coalesced) against x_R staged in LDS.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_bwd_rpart(SolveArgs<TF> a, const int32_t* __restrict__ tasks,
                                                   const int64_t* __restrict__ part_off, const TX* __restrict__ x,
                                                   TX* __restrict__ part) {
  __shared__ TX xs[kRpartRows];
  const int s = tasks[3 * blockIdx.x], j = tasks[3 * blockIdx.x + 1], ch = tasks[3 * blockIdx.x + 2];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int c0 = j * kRB;
  const int bw = min(kRB, nc - c0);
  const int nch = (nr + kRpartRows - 1) / kRpartRows;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const TF* P = a.panel + a.panel_off[s];
  const int32_t* R = a.rows + a.rows_off[s];
  const int q0 = ch * kRpartRows, len = min(kRpartRows, nr - q0);
  for (int q = tid; q < len; q += blockDim.x) xs[q] = x[R[q0 + q]];
  __syncthreads();
  TX* out = part + part_off[s] + static_cast<int64_t>(j) * nch * kRB + static_cast<int64_t>(ch) * kRB;
  // wave w: columns w, w + 4, ... of the block, four at a time
  for (int t0 = w * 4; t0 < bw; t0 += 16) {
    TX acc[4] = {TX(0), TX(0), TX(0), TX(0)};
    const TF* col[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      col[u] = P + (nc + q0) + static_cast<int64_t>(c0 + min(t0 + u, bw - 1)) * m;
    for (int q = lane; q < len; q += 64) {
      const TX xv = xs[q];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += static_cast<TX>(col[u][q]) * xv;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      TX v = acc[u];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0 && t0 + u < bw) out[t0 + u] = v;
    }
  }
}

// Backward, column block j of front s: z = x(block) - sum over the rows below
// of L[r, block]^T x_r (R rows are final; rows of later column blocks wait
// for those blocks' flags), then the unit-upper solve with L_jj^T. The tile
// of the last dependency (block j+1) is staged transposed in LDS before the
// wait, so after its flag only 64 LDS dot products, the diagonal solve and
// the publish remain.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_bwd_front(SolveArgs<TF> a, const int32_t* __restrict__ tasks,
                                                   const int32_t* __restrict__ flag_off, int* flags, int epoch,
                                                   TX* x, const int64_t* __restrict__ part_off,
                                                   const TX* __restrict__ part, int* err) {
  __shared__ TX red[4][kRB];
  __shared__ TX xs[kRB];
  __shared__ TF Ld[kRB][kRB + 1];
  __shared__ TF Tt[kRB][kRB + 1];  // critical tile, Tt[t][r] = L[rb0 + r, c0 + t]
  const int s = tasks[2 * blockIdx.x], j = tasks[2 * blockIdx.x + 1];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int c0 = j * kRB;
  const int bw = min(kRB, nc - c0);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const TF* P = a.panel + a.panel_off[s];
  const int32_t* R = a.rows + a.rows_off[s];
  const int ncb = (nc + kRB - 1) / kRB;
  const int rb0 = c0 + bw;                 // first row below the block
  const bool crit = j + 1 < ncb;           // tile 0 = rows [rb0, rb0 + 64) holds block j+1
  const int crow = crit ? min(kRB, nc - rb0) : 0;  // its rows (columns of s only)
  for (int t = w; t < bw; t += 4) {
    if (lane < bw) Ld[lane][t] = P[(c0 + lane) + static_cast<int64_t>(c0 + t) * m];
    if (lane < crow) Tt[t][lane] = P[(rb0 + lane) + static_cast<int64_t>(c0 + t) * m];
  }
  TX acc[kRB];
#pragma unroll
  for (int t = 0; t < kRB; ++t) acc[t] = TX(0);
  // rows of later column blocks in 64-row tiles, oldest block first; the
  // critical tile (q = 0 when crit) is left to the LDS path, and the R rows
  // after the last column block come from the partial sums
  const int q_first = crit ? 1 : 0;
  const int ntiles = crit ? (nc - rb0 + kRB - 1) / kRB : 0;
  for (int q = ntiles - 1 - w; q >= q_first; q -= 4) {
    const int rs = rb0 + q * kRB;
    const int r = rs + lane;
    const int rc = r < nc ? r : rs;
    const TF* row = P + rc + static_cast<int64_t>(c0) * m;
    TF lv[kRB];
#pragma unroll
    for (int t = 0; t < kRB; ++t) lv[t] = t < bw ? row[static_cast<int64_t>(t) * m] : TF(0);
    if (lane == 0) wait_flag(flags + flag_off[s] + rs / kRB, epoch, err);
    const TX xr = r < nc ? ld_dev(x + f + r) : TX(0);
#pragma unroll
    for (int t = 0; t < kRB; ++t) acc[t] += static_cast<TX>(lv[t]) * xr;
  }
#pragma unroll
  for (int t = 0; t < kRB; ++t) {
    TX v = acc[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == t) red[w][t] = v;
  }
  __syncthreads();
  if (w == 0) {
    TX z = lane < bw ? x[f + c0 + lane] - ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane])) : TX(0);
    // R rows beyond the last column block's tile: partial sums in fixed order
    const int nch = (nr + kRpartRows - 1) / kRpartRows;
    if (lane < bw) {
      const TX* pp = part + part_off[s] + static_cast<int64_t>(j) * nch * kRB + lane;
      TX rs = TX(0);
      for (int c = 0; c < nch; ++c) rs += pp[static_cast<int64_t>(c) * kRB];
      z -= rs;
    }
    if (crit) {
      if (lane == 0) wait_flag(flags + flag_off[s] + j + 1, epoch, err);
      xs[lane] = lane < crow ? ld_dev(x + f + rb0 + lane) : TX(0);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      if (lane < bw) {
        TX cs = TX(0);
        for (int r = 0; r < crow; ++r) cs += static_cast<TX>(Tt[lane][r]) * xs[r];
        z -= cs;
      }
    }
    for (int t = bw - 1; t >= 0; --t) {
      const TX xt = __shfl(z, t);
      if (lane < t) z -= static_cast<TX>(Ld[t][lane]) * xt;
    }
    if (lane < bw) st_dev(x + f + c0 + lane, z);
  }
  publish_flag(flags + flag_off[s] + j, epoch);
}

}  // namespace

template <class TF, class TX>
void ldl_solve(const SolveArgs<TF>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, int epoch, hipStream_t st) {
  TX* ybuf = reinterpret_cast<TX*>(dp.ybuf);
  // forward, leaves to root; D^{-1} fused
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_fwd_small<TF, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x, uvec);
    const int t0 = dp.h_fwd_ptr[l], t1 = dp.h_fwd_ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL((k_fwd_front<TF, TX>), dim3(t1 - t0), dim3(256), 0, st, a, dp.fwd_task + 2 * t0,
                         dp.fwd_child_ptr + t0, dp.fwd_child, dp.flag_off, dp.flags, 2 * epoch + 1, x, uvec, ybuf,
                         dp.err);
  }
  // backward, root to leaves
  TX* part = reinterpret_cast<TX*>(dp.part);
  for (int l = dp.nlevels - 1; l >= 0; --l) {
    const int p0 = dp.h_rpart_ptr[l], p1 = dp.h_rpart_ptr[l + 1];
    if (p1 > p0)
      hipLaunchKernelGGL((k_bwd_rpart<TF, TX>), dim3(p1 - p0), dim3(256), 0, st, a, dp.rpart_task + 3 * p0,
                         dp.part_off, x, part);
    const int t0 = dp.h_bwd_ptr[l], t1 = dp.h_bwd_ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL((k_bwd_front<TF, TX>), dim3(t1 - t0), dim3(256), 0, st, a, dp.bwd_task + 2 * t0, dp.flag_off,
                         dp.flags, 2 * epoch + 2, x, dp.part_off, part, dp.err);
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0) hipLaunchKernelGGL((k_bwd_small<TF, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x);
  }
  MPIR_HIP(hipGetLastError());
}

template void ldl_solve<double, double>(const SolveArgs<double>&, const DeviceSolvePlan&, double*, double*, int, hipStream_t);
template void ldl_solve<float, double>(const SolveArgs<float>&, const DeviceSolvePlan&, double*, double*, int, hipStream_t);
template void ldl_solve<float, float>(const SolveArgs<float>&, const DeviceSolvePlan&, float*, float*, int, hipStream_t);

}  // namespace mpir
