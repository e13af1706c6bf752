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
//
// Big fronts hold Z = L_11^{-T} D^{-1} (upper triangular) below their panel,
// so their solves are matrix-vector products without a dependency chain:
//   forward   x_c = Z^T v_c  (= D^{-1} L_11^{-1} v_c),  u = v_R - L_21 D x_c
//   backward  x_c = Z D (x_c - L_21^T x_R)          (= L_11^{-T} (...))
// Every sum runs in a fixed order, so repeated solves agree bit for bit.
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
constexpr int kRB = kSolveRB;
constexpr int kChunk = kRpartRows;  // vector entries staged in LDS at a time

// Forward gather, 64-row block i of front s: v = [b on columns; 0 on R] +
// the children's update vectors, children in fixed order. The column part
// is written back into x, the R part into the front's update vector.
template <class TX>
__global__ void __launch_bounds__(64) k_fwd_gather(const int32_t* __restrict__ first, const int32_t* __restrict__ nrow,
                                                   const int64_t* __restrict__ rows_off,
                                                   const int32_t* __restrict__ relidx, const int32_t* __restrict__ tasks,
                                                   const int32_t* __restrict__ cptr, const int32_t* __restrict__ cl,
                                                   TX* x, TX* uvec) {
  __shared__ TX vsh[kRB];
  const int s = tasks[2 * blockIdx.x], i = tasks[2 * blockIdx.x + 1];
  const int f = first[s];
  const int nc = first[s + 1] - f;
  const int m = nc + nrow[s];
  const int r0 = i * kRB;
  const int lane = threadIdx.x;
  const int r = r0 + lane;
  vsh[lane] = r < nc ? x[f + r] : TX(0);
  __syncthreads();
  for (int e = cptr[blockIdx.x]; e < cptr[blockIdx.x + 1]; ++e) {
    const int c = cl[3 * e], q0 = cl[3 * e + 1], q1 = cl[3 * e + 2];
    const int q = q0 + lane;
    if (q < q1) vsh[relidx[rows_off[c] + q] - r0] += uvec[rows_off[c] + q];
    __syncthreads();
  }
  if (r < nc) x[f + r] = vsh[lane];
  else if (r < m) uvec[rows_off[s] + (r - nc)] = vsh[lane];
}

// Forward, column block j of front s: xf[col] = sum_{i <= col} Z[i, col] v_i
// over the block's columns (Z is zero below its diagonal, so whole 64-row
// tiles are summed). Columns of Z are contiguous: each wave walks four
// columns at a time down the rows, v staged in LDS.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_fwd_xc(SolveArgs<TF> a, const int32_t* __restrict__ tasks,
                                                const TX* __restrict__ x, TX* __restrict__ xf) {
  __shared__ TX vs[kChunk];
  const int s = tasks[2 * blockIdx.x], j = tasks[2 * blockIdx.x + 1];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int c0 = j * kRB;
  const int bw = min(kRB, nc - c0);
  const int rend = c0 + bw;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const TF* Z = a.panel + a.panel_off[s] + m;
  TX acc[4][4];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[g][u] = TX(0);
  for (int q0 = 0; q0 < rend; q0 += kChunk) {
    const int len = min(kChunk, rend - q0);
    __syncthreads();
    for (int q = tid; q < len; q += blockDim.x) vs[q] = x[f + q0 + q];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int t0 = g * 16 + w * 4;
      if (t0 >= bw) break;
      const TF* col[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) col[u] = Z + q0 + (c0 + min(t0 + u, bw - 1)) * ld;
      for (int q = lane; q < len; q += 64) {
        const TX v = vs[q];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[g][u] += static_cast<TX>(col[u][q]) * v;
      }
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int t0 = g * 16 + w * 4;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      TX v = acc[g][u];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0 && t0 + u < bw) xf[f + c0 + t0 + u] = v;
    }
  }
}

// Row-block GEMV over a panel: out[r] = sum_{c in [cb, ce)} M[r, c] y_c for
// the 64 rows r0.. of the block, y_c = scale[c] * (yv[c] - sum of nch
// partial sums) staged in LDS. Each lane owns a row, the four waves split
// the columns, and the wave results are added in a fixed order.
template <class TF, class TX>
__device__ __forceinline__ TX rowblock_gemv(const TF* __restrict__ M, int64_t ld, int rows, int cb, int ce,
                                            const TF* __restrict__ scale, const TX* __restrict__ yv,
                                            const TX* __restrict__ part, int nch, TX* ys, TX (*red)[kRB]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rr = lane < rows ? lane : 0;
  TX acc = TX(0);
  for (int q0 = cb; q0 < ce; q0 += kChunk) {
    const int len = min(kChunk, ce - q0);
    __syncthreads();
    for (int q = tid; q < len; q += blockDim.x) {
      const int c = q0 + q;
      TX z = yv[c];
      if (part != nullptr) {
        const TX* pp = part + static_cast<int64_t>(c / kRB) * nch * kRB + (c % kRB);
        TX ps = TX(0);
        for (int h = 0; h < nch; ++h) ps += pp[static_cast<int64_t>(h) * kRB];
        z -= ps;
      }
      ys[q] = static_cast<TX>(scale[c]) * z;
    }
    __syncthreads();
    // wave w: columns [w * per, (w + 1) * per) of the chunk, eight at a time
    const int per = (len + 3) / 4;
    const int cs = min(len, w * per), cend = min(len, cs + per);
    const TF* mp = M + rr + (q0 + cs) * ld;
    int c = cs;
    for (; c + 8 <= cend; c += 8, mp += 8 * ld) {
      TF lv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) lv[u] = mp[u * ld];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += static_cast<TX>(lv[u]) * ys[c + u];
    }
    for (; c < cend; ++c, mp += ld) acc += static_cast<TX>(*mp) * ys[c];
  }
  red[w][lane] = acc;
  __syncthreads();
  return (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// Forward, 64-row block b of R_s: u = v_R - L_21 D xf.
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_fwd_u(SolveArgs<TF> a, const int32_t* __restrict__ tasks,
                                               const TX* __restrict__ xf, TX* __restrict__ uvec) {
  __shared__ TX ys[kChunk];
  __shared__ TX red[4][kRB];
  const int s = tasks[2 * blockIdx.x], b = tasks[2 * blockIdx.x + 1];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int64_t ld = panel_ld(nc, nr);
  const int rb0 = b * kRB;
  const int rows = min(kRB, nr - rb0);
  const TF* M = a.panel + a.panel_off[s] + nc + rb0;
  const TX sum = rowblock_gemv<TF, TX>(M, ld, rows, 0, nc, a.D + f, xf + f, static_cast<const TX*>(nullptr), 0, ys, red);
  const int lane = threadIdx.x;
  if (lane < rows) {
    TX* u = uvec + a.rows_off[s] + rb0;
    u[lane] = u[lane] - sum;
  }
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
  const int64_t ld = panel_ld(nc, nr);
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
  // wave w: columns w * 4, w * 4 + 16, ... of the block, four at a time
  for (int t0 = w * 4; t0 < bw; t0 += 16) {
    TX acc[4] = {TX(0), TX(0), TX(0), TX(0)};
    const TF* col[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) col[u] = P + (nc + q0) + (c0 + min(t0 + u, bw - 1)) * ld;
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

// Backward, 64-row block i of the columns of front s:
// x_r = sum_{c >= 64 i} Z[r, c] D_c z_c, z = xf - (partial sums of L_21^T x_R).
template <class TF, class TX>
__global__ void __launch_bounds__(256) k_bwd_xc(SolveArgs<TF> a, const int32_t* __restrict__ tasks,
                                                const int64_t* __restrict__ part_off, const TX* __restrict__ part,
                                                const TX* __restrict__ xf, TX* __restrict__ x) {
  __shared__ TX ys[kChunk];
  __shared__ TX red[4][kRB];
  const int s = tasks[2 * blockIdx.x], i = tasks[2 * blockIdx.x + 1];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int r0 = i * kRB;
  const int rows = min(kRB, nc - r0);
  const int nch = (nr + kRpartRows - 1) / kRpartRows;
  const TF* M = a.panel + a.panel_off[s] + m + r0;
  // the partial sums are laid out by column block: pass them shifted so that
  // column c (absolute within the front) finds block c / 64
  const TX* pp = nch > 0 ? part + part_off[s] : static_cast<const TX*>(nullptr);
  const TX sum = rowblock_gemv<TF, TX>(M, ld, rows, r0, nc, a.D + f, xf + f, pp, nch, ys, red);
  const int lane = threadIdx.x;
  if (lane < rows) x[f + r0 + lane] = sum;
}

}  // namespace

template <class TF, class TX>
void ldl_solve(const SolveArgs<TF>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st) {
  TX* xf = reinterpret_cast<TX*>(dp.xf);
  TX* part = reinterpret_cast<TX*>(dp.part);
  // forward, leaves to root; D^{-1} fused
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_fwd_small<TF, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x, uvec);
    const int g0 = dp.h_gather_ptr[l], g1 = dp.h_gather_ptr[l + 1];
    if (g1 > g0)
      hipLaunchKernelGGL((k_fwd_gather<TX>), dim3(g1 - g0), dim3(64), 0, st, a.first, a.nrow, a.rows_off, a.relidx,
                         dp.gather_task + 2 * g0, dp.gather_child_ptr + g0, dp.gather_child, x, uvec);
    const int c0 = dp.h_xc_ptr[l], c1 = dp.h_xc_ptr[l + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_fwd_xc<TF, TX>), dim3(c1 - c0), dim3(256), 0, st, a, dp.xc_task + 2 * c0, x, xf);
    const int u0 = dp.h_u_ptr[l], u1 = dp.h_u_ptr[l + 1];
    if (u1 > u0)
      hipLaunchKernelGGL((k_fwd_u<TF, TX>), dim3(u1 - u0), dim3(256), 0, st, a, dp.u_task + 2 * u0, xf, uvec);
  }
  // backward, root to leaves
  for (int l = dp.nlevels - 1; l >= 0; --l) {
    const int p0 = dp.h_rpart_ptr[l], p1 = dp.h_rpart_ptr[l + 1];
    if (p1 > p0)
      hipLaunchKernelGGL((k_bwd_rpart<TF, TX>), dim3(p1 - p0), dim3(256), 0, st, a, dp.rpart_task + 3 * p0,
                         dp.part_off, x, part);
    const int t0 = dp.h_bwd_ptr[l], t1 = dp.h_bwd_ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL((k_bwd_xc<TF, TX>), dim3(t1 - t0), dim3(256), 0, st, a, dp.bwd_task + 2 * t0, dp.part_off,
                         part, xf, x);
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0) hipLaunchKernelGGL((k_bwd_small<TF, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x);
  }
  MPIR_HIP(hipGetLastError());
}

template void ldl_solve<double, double>(const SolveArgs<double>&, const DeviceSolvePlan&, double*, double*, hipStream_t);
template void ldl_solve<float, double>(const SolveArgs<float>&, const DeviceSolvePlan&, double*, double*, hipStream_t);
template void ldl_solve<float, float>(const SolveArgs<float>&, const DeviceSolvePlan&, float*, float*, hipStream_t);

}  // namespace mpir
