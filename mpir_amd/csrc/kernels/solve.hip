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
// Big fronts hold W = L_11^{-T} (unit upper triangular) below their panel,
// so their solves are matrix-vector products without a dependency chain:
//   forward   y_c = W^T v_c  (= L_11^{-1} v_c),  u = v_R - L_21 y_c
//   backward  x_c = W (D^{-1} y_c - L_21^T x_R)  (= L_11^{-T} (...))
// Every sum runs in a fixed order, so repeated solves agree bit for bit.
// The factor is stored in TF (double, float, half or bfloat16), D^{-1} in
// TD (TF, or float for the 16-bit factors); products and sums run in TX.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common.hpp"
#include "kernels.hpp"

namespace mpir {

namespace {

// ---------------------------------------------------------------- small fronts
template <class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_fwd_small(SolveArgs<TF, TD> a, const int32_t* __restrict__ list, TX* __restrict__ x,
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

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_bwd_small(SolveArgs<TF, TD> a, const int32_t* __restrict__ list, TX* __restrict__ x) {
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
constexpr int kChunk = 2048;  // vector entries staged in LDS at a time (row-block GEMV)

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

// Dot products of the bw (<= 64) columns base + t * ld, rows [0, len), with
// vs (LDS): out[t] = sum_q base[q + t * ld] vs[q]. Wave w takes columns
// 4w .. 4w + 3, then 16 further, lanes stride the rows (coalesced); four
// row strides are loaded before use so each wave keeps 16 loads in flight.
// Partial sums are combined across lanes in a fixed order.
template <class TF, class TX>
__device__ __forceinline__ void column_dots(const TF* __restrict__ base, int64_t ld, int bw, int len,
                                            const TX* __restrict__ vs, TX* __restrict__ out) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int t0 = w * 4; t0 < bw; t0 += 16) {
    const TF* col[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) col[u] = base + min(t0 + u, bw - 1) * ld;
    TX acc[4] = {TX(0), TX(0), TX(0), TX(0)};
    int q = lane;
    for (; q + 192 < len; q += 256) {
      TF lv[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int u = 0; u < 4; ++u) lv[r][u] = col[u][q + 64 * r];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const TX v = vs[q + 64 * r];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += static_cast<TX>(lv[r][u]) * v;
      }
    }
    for (; q < len; q += 64) {
      const TX v = vs[q];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += static_cast<TX>(col[u][q]) * v;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      TX v = acc[u];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0 && t0 + u < bw) out[t0 + u] = v;
    }
  }
}

// Forward, rows [h * kXcRows, ...) of column block j of front s: partial sums
// of y[col] = sum_{i <= col} W[i, col] v_i over the block's columns (W is
// zero below its diagonal, so whole 64-row tiles are summed).
template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_fwd_xc(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                                const int64_t* __restrict__ xfp_off, const TX* __restrict__ x,
                                                TX* __restrict__ xfp) {
  __shared__ TX vs[kXcRows];
  const int s = tasks[3 * blockIdx.x], j = tasks[3 * blockIdx.x + 1], h = tasks[3 * blockIdx.x + 2];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int c0 = j * kRB;
  const int bw = min(kRB, nc - c0);
  const int q0 = h * kXcRows;
  const int len = min(kXcRows, c0 + bw - q0);
  const int nhx = (nc + kXcRows - 1) / kXcRows;
  for (int q = threadIdx.x; q < len; q += blockDim.x) vs[q] = x[f + q0 + q];
  __syncthreads();
  const TF* base = a.panel + a.panel_off[s] + m + q0 + c0 * ld;
  column_dots<TF, TX>(base, ld, bw, len, vs, xfp + xfp_off[s] + (static_cast<int64_t>(j) * nhx + h) * kRB);
}

// y_c of column c of a big front after the forward pass: the sum of its
// partials.
template <class TX>
__device__ __forceinline__ TX forward_value(const TX* __restrict__ xfp_s, int nc, int c) {
  const int nhx = (nc + kXcRows - 1) / kXcRows;
  const int j = c / kRB;
  const int nv = (min(nc, (j + 1) * kRB) + kXcRows - 1) / kXcRows;
  const TX* pp = xfp_s + static_cast<int64_t>(j) * nhx * kRB + (c % kRB);
  TX v = TX(0);
  for (int h = 0; h < nv; ++h) v += pp[static_cast<int64_t>(h) * kRB];
  return v;
}

// Row-block GEMV over a panel: returns sum_{c in [cb, ce)} M[r, c] y_c for the
// row r = lane (< rows) of the 64-row block at M, with y_c = stage(c) staged
// in LDS. Each lane owns a row, the four waves split the columns (16 loads
// in flight each), and the wave results are added in a fixed order.
template <class TF, class TX, class Stage>
__device__ __forceinline__ TX rowblock_gemv(const TF* __restrict__ M, int64_t ld, int rows, int cb, int ce,
                                            Stage stage, TX* ys, TX (*red)[kRB]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rr = lane < rows ? lane : 0;
  TX acc = TX(0);
  for (int q0 = cb; q0 < ce; q0 += kChunk) {
    const int len = min(kChunk, ce - q0);
    __syncthreads();
    for (int q = tid; q < len; q += blockDim.x) ys[q] = stage(q0 + q);
    __syncthreads();
    // wave w: columns [w * per, (w + 1) * per) of the chunk, 16 at a time
    const int per = (len + 3) / 4;
    const int cs = min(len, w * per), cend = min(len, cs + per);
    const TF* mp = M + rr + (q0 + cs) * ld;
    int c = cs;
    for (; c + 16 <= cend; c += 16, mp += 16 * ld) {
      TF lv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) lv[u] = mp[u * ld];
#pragma unroll
      for (int u = 0; u < 16; ++u) acc += static_cast<TX>(lv[u]) * ys[c + u];
    }
    for (; c < cend; ++c, mp += ld) acc += static_cast<TX>(*mp) * ys[c];
  }
  red[w][lane] = acc;
  __syncthreads();
  return (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// Forward, 64-row block b of R_s: u = v_R - L_21 y.
template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_fwd_u(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                               const int64_t* __restrict__ xfp_off, const TX* __restrict__ xfp,
                                               TX* __restrict__ uvec) {
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
  const TX* xs = xfp + xfp_off[s];
  const TX sum = rowblock_gemv<TF, TX>(M, ld, rows, 0, nc, [=](int c) { return forward_value(xs, nc, c); }, ys, red);
  const int lane = threadIdx.x;
  if (lane < rows) {
    TX* u = uvec + a.rows_off[s] + rb0;
    u[lane] = u[lane] - sum;
  }
}

// Backward, rows below the columns of a front (R rows, all final): partial
// sums part[(j * nch + c) * 64 + t] = sum over the chunk's rows r of
// L[nc + r, 64 j + t] x[R_r].
template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_bwd_rpart(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
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
  const int32_t* R = a.rows + a.rows_off[s];
  const int q0 = ch * kRpartRows, len = min(kRpartRows, nr - q0);
  for (int q = threadIdx.x; q < len; q += blockDim.x) xs[q] = x[R[q0 + q]];
  __syncthreads();
  const TF* base = a.panel + a.panel_off[s] + nc + q0 + c0 * ld;
  column_dots<TF, TX>(base, ld, bw, len, xs,
                      part + part_off[s] + static_cast<int64_t>(j) * nch * kRB + static_cast<int64_t>(ch) * kRB);
}

// Backward, 64-row block i of the columns of front s:
// x_r = sum_{c >= 64 i} W[r, c] z_c, z = D^{-1} y - (partial sums of L_21^T x_R).
template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_bwd_xc(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                                const int64_t* __restrict__ part_off, const TX* __restrict__ part,
                                                const int64_t* __restrict__ xfp_off, const TX* __restrict__ xfp,
                                                TX* __restrict__ x) {
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
  const TX* xs = xfp + xfp_off[s];
  const TX* ps = part + (nch > 0 ? part_off[s] : 0);
  const TD* Dinv = a.Dinv + f;
  const TX sum = rowblock_gemv<TF, TX>(
      M, ld, rows, r0, nc,
      [=](int c) {
        const TX z = static_cast<TX>(Dinv[c]) * forward_value(xs, nc, c);
        const TX* pp = ps + static_cast<int64_t>(c / kRB) * nch * kRB + (c % kRB);
        TX rs = TX(0);
        for (int h = 0; h < nch; ++h) rs += pp[static_cast<int64_t>(h) * kRB];
        return z - rs;
      },
      ys, red);
  const int lane = threadIdx.x;
  if (lane < rows) x[f + r0 + lane] = sum;
}

}  // namespace

template <class TF, class TD, class TX>
void ldl_solve(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st) {
  TX* xfp = reinterpret_cast<TX*>(dp.xfp);
  TX* part = reinterpret_cast<TX*>(dp.part);
  // forward, leaves to root; D^{-1} fused
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_fwd_small<TF, TD, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x, uvec);
    const int g0 = dp.h_gather_ptr[l], g1 = dp.h_gather_ptr[l + 1];
    if (g1 > g0)
      hipLaunchKernelGGL((k_fwd_gather<TX>), dim3(g1 - g0), dim3(64), 0, st, a.first, a.nrow, a.rows_off, a.relidx,
                         dp.gather_task + 2 * g0, dp.gather_child_ptr + g0, dp.gather_child, x, uvec);
    const int c0 = dp.h_xc_ptr[l], c1 = dp.h_xc_ptr[l + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_fwd_xc<TF, TD, TX>), dim3(c1 - c0), dim3(256), 0, st, a, dp.xc_task + 3 * c0, dp.xfp_off, x,
                         xfp);
    const int u0 = dp.h_u_ptr[l], u1 = dp.h_u_ptr[l + 1];
    if (u1 > u0)
      hipLaunchKernelGGL((k_fwd_u<TF, TD, TX>), dim3(u1 - u0), dim3(256), 0, st, a, dp.u_task + 2 * u0, dp.xfp_off, xfp,
                         uvec);
  }
  // backward, root to leaves
  for (int l = dp.nlevels - 1; l >= 0; --l) {
    const int p0 = dp.h_rpart_ptr[l], p1 = dp.h_rpart_ptr[l + 1];
    if (p1 > p0)
      hipLaunchKernelGGL((k_bwd_rpart<TF, TD, TX>), dim3(p1 - p0), dim3(256), 0, st, a, dp.rpart_task + 3 * p0,
                         dp.part_off, x, part);
    const int t0 = dp.h_bwd_ptr[l], t1 = dp.h_bwd_ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL((k_bwd_xc<TF, TD, TX>), dim3(t1 - t0), dim3(256), 0, st, a, dp.bwd_task + 2 * t0, dp.part_off,
                         part, dp.xfp_off, xfp, x);
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0) hipLaunchKernelGGL((k_bwd_small<TF, TD, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x);
  }
  MPIR_HIP(hipGetLastError());
}

template void ldl_solve<double, double, double>(const SolveArgs<double, double>&, const DeviceSolvePlan&, double*,
                                                 double*, hipStream_t);
template void ldl_solve<float, float, double>(const SolveArgs<float, float>&, const DeviceSolvePlan&, double*, double*,
                                              hipStream_t);
template void ldl_solve<float, float, float>(const SolveArgs<float, float>&, const DeviceSolvePlan&, float*, float*,
                                             hipStream_t);
template void ldl_solve<__half, float, double>(const SolveArgs<__half, float>&, const DeviceSolvePlan&, double*,
                                               double*, hipStream_t);
template void ldl_solve<__hip_bfloat16, float, double>(const SolveArgs<__hip_bfloat16, float>&,
                                                       const DeviceSolvePlan&, double*, double*, hipStream_t);

}  // namespace mpir
