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
// transposed.
//
// Big fronts hold W = L_11^{-T} (unit upper triangular) below their panel,
// so their solves are matrix-vector products without a dependency chain:
//   forward   y_c = W^T v_c  (= L_11^{-1} v_c),  u = v_R - L_21 y_c
//   backward  x_c = W (D^{-1} y_c - L_21^T x_R)  (= L_11^{-T} (...))
// Every sum runs in a fixed order, so repeated solves agree bit for bit.
// The factor is stored in TF (double, float, half or bfloat16), D^{-1} in
// TD (TF, or float for the 16-bit factors); products and sums run in TX.
//
// Two schedules run the same task bodies, so their results agree bit for bit:
//   levels: one launch per task kind and level of the supernodal tree;
//   dag:    one persistent launch per direction. Workgroups take tasks in a
//           topological order from a dispenser and wait (bounded) on
//           per-front completion counters, so a front starts as soon as its
//           own children (forward) or its parent (backward) are done. Values
//           passed between workgroups of a launch use device-coherent
//           stores and loads.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../common.hpp"
#include "kernels.hpp"

namespace mpir {

namespace {

constexpr int kRB = kSolveRB;
constexpr int kChunk = 2048;  // vector entries staged in LDS at a time (row-block GEMV)
constexpr int kSpinLimit = 1 << 20;  // polls (about a second) before a wait gives up
// LDS of one 256-thread solve task: the GEMV stage and its wave reduction
constexpr int kSmem = kChunk + 4 * kRB;
static_assert(kXcRows <= kChunk && kRpartRows <= kChunk, "staging buffers share the GEMV stage");

// ---------------------------------------------------------------- memory
// Relaxed device-scope atomic accesses: plain loads and stores with the
// device-coherence bit, no cache maintenance. C selects them for values
// that another workgroup of the same launch writes or reads.
template <class T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool C, class T>
__device__ __forceinline__ T ldx(const T* p) {
  if constexpr (C) return ld_dev(p);
  else return *p;
}
template <bool C, class T>
__device__ __forceinline__ void stx(T* p, T v) {
  if constexpr (C) st_dev(p, v);
  else *p = v;
}

// orders the LDS accesses of one wavefront (no other wave takes part)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of v in lane j (j uniform across the wave)
template <class T>
__device__ __forceinline__ T lane_value(T v, int j) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(b), j);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(b >> 32), j);
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), j));
  }
}

struct NoWait {
  __device__ void operator()() const {}
};

// ---------------------------------------------------------------- small fronts
// One wavefront per front (m <= 64), lane = row; v: 64 entries of LDS. L is
// read kLC columns (rows) at a time, all loads of a group in flight at once.
// wait() runs after the front's first loads and before anything another
// task writes is read.
constexpr int kLC = 16;

template <bool C, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void small_fwd(const SolveArgs<TF, TD>& a, int s, TX* x, TX* uvec, TX* v, Wait wait) {
  const int lane = threadIdx.x & 63;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const TF* P = a.panel + a.panel_off[s];
  // this lane's row of L, columns [j0, j0 + kLC) clamped into the panel
  const TF* prow = P + min(lane, m - 1);
  TF lr[kLC];
  auto load = [&](int j0) {
#pragma unroll
    for (int u = 0; u < kLC; ++u) lr[u] = prow[static_cast<int64_t>(min(j0 + u, nc - 1)) * m];
  };
  load(0);
  const TX b = x[f + min(lane, nc - 1)];
  const TD dinv = a.Dinv[f + min(lane, nc - 1)];
  v[lane] = lane < nc ? b : TX(0);
  wait();
  for (int t = a.child_ptr[s]; t < a.child_ptr[s + 1]; ++t) {
    const int c = a.child_list[t];
    const int nrc = a.nrow[c];
    if (lane < nrc) {
      const int64_t q = a.rows_off[c] + lane;
      v[a.relidx[q]] += ldx<C>(uvec + q);
    }
    wave_sync();
  }
  TX vl = v[lane];
  for (int j0 = 0; j0 < nc; j0 += kLC) {
    if (j0 > 0) load(j0);
#pragma unroll
    for (int u = 0; u < kLC; ++u) {
      const int j = j0 + u;
      if (j < nc) {
        const TX yj = lane_value(vl, j);
        if (lane > j) vl -= static_cast<TX>(lr[u]) * yj;
      }
    }
  }
  if (lane < nc) x[f + lane] = vl * static_cast<TX>(dinv);
  else if (lane < m) stx<C>(uvec + a.rows_off[s] + (lane - nc), vl);
}

template <bool C, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void small_bwd(const SolveArgs<TF, TD>& a, int s, TX* x, TX* w, Wait wait) {
  const int lane = threadIdx.x & 63;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const TF* P = a.panel + a.panel_off[s];
  // column `lane` of L, rows clamped into the panel
  const int cl = min(lane, nc - 1);
  const TF* pcol = P + static_cast<int64_t>(cl) * m;
  TF lc[kLC];
  auto load_up = [&](int r0) {  // rows r0, r0 + 1, ...
#pragma unroll
    for (int u = 0; u < kLC; ++u) lc[u] = pcol[min(r0 + u, m - 1)];
  };
  auto load_down = [&](int t1) {  // rows t1 - 1, t1 - 2, ...
#pragma unroll
    for (int u = 0; u < kLC; ++u) lc[u] = pcol[max(t1 - 1 - u, 0)];
  };
  load_up(nc);
  const int32_t row = nr > 0 ? a.rows[a.rows_off[s] + min(max(lane - nc, 0), nr - 1)] : 0;
  const TX own = x[f + cl];
  wait();
  if (lane < nc) w[lane] = own;
  else if (lane < m) w[lane] = ldx<C>(x + row);
  wave_sync();
  // lane j < nc: z_j = w_j - sum over rows r >= nc of L[r, j] w_r, r ascending
  TX z = own;
  for (int r0 = nc; r0 < m; r0 += kLC) {
    if (r0 > nc) load_up(r0);
#pragma unroll
    for (int u = 0; u < kLC; ++u)
      if (r0 + u < m) z -= static_cast<TX>(lc[u]) * w[r0 + u];
  }
  // unit upper solve with L_11^T, bottom to top
  for (int t1 = nc; t1 > 0; t1 -= kLC) {
    load_down(t1);
#pragma unroll
    for (int u = 0; u < kLC; ++u) {
      const int t = t1 - 1 - u;
      if (t >= 0) {
        const TX xt = lane_value(z, t);
        if (lane < t) z -= static_cast<TX>(lc[u]) * xt;
      }
    }
  }
  if (lane < nc) stx<C>(x + f + lane, z);
}

// ---------------------------------------------------------------- big fronts
// Forward gather, 64-row block i of front s (one wavefront): v = [b on
// columns; 0 on R] + the children's update vectors (entries [e0, e1) of cl),
// children in fixed order; the column part goes back into x, the R part into
// the front's update vector.
template <bool C, class TF, class TD, class TX>
__device__ __forceinline__ void gather_rows(const SolveArgs<TF, TD>& a, int s, int i, const int32_t* __restrict__ cl,
                                            int e0, int e1, TX* x, TX* uvec, TX* vsh) {
  const int lane = threadIdx.x & 63;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int r0 = i * kRB;
  const int r = r0 + lane;
  vsh[lane] = r < nc ? x[f + r] : TX(0);
  wave_sync();
  for (int e = e0; e < e1; ++e) {
    const int c = cl[3 * e], q0 = cl[3 * e + 1], q1 = cl[3 * e + 2];
    const int q = q0 + lane;
    if (q < q1) {
      const int64_t o = a.rows_off[c] + q;
      vsh[a.relidx[o] - r0] += ldx<C>(uvec + o);
    }
    wave_sync();
  }
  if (r < nc) stx<C>(x + f + r, vsh[lane]);
  else if (r < m) stx<C>(uvec + a.rows_off[s] + (r - nc), vsh[lane]);
}

// Dot products of the bw (<= 64) columns base + t * ld, rows [0, len), with
// vs (LDS): out[t] = sum_q base[q + t * ld] vs[q]. Wave w takes columns
// 4w .. 4w + 3, then 16 further, lanes stride the rows (coalesced); four
// row strides are loaded before use so each wave keeps 16 loads in flight.
// Partial sums are combined across lanes in a fixed order.
template <bool C, class TF, class TX>
__device__ __forceinline__ void column_dots(const TF* __restrict__ base, int64_t ld, int bw, int len,
                                            const TX* __restrict__ vs, TX* out) {
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
      if (lane == 0 && t0 + u < bw) stx<C>(out + t0 + u, v);
    }
  }
}

// Forward, rows [h * kXcRows, ...) of column block j of front s: partial sums
// of y[col] = sum_{i <= col} W[i, col] v_i over the block's columns (W is
// zero below its diagonal, so whole 64-row tiles are summed).
template <bool C, class TF, class TD, class TX>
__device__ __forceinline__ void fwd_xc(const SolveArgs<TF, TD>& a, int s, int j, int h,
                                       const int64_t* __restrict__ xfp_off, const TX* x, TX* xfp, TX* vs) {
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
  for (int q = threadIdx.x; q < len; q += blockDim.x) vs[q] = ldx<C>(x + f + q0 + q);
  __syncthreads();
  const TF* base = a.panel + a.panel_off[s] + m + q0 + c0 * ld;
  column_dots<C, TF, TX>(base, ld, bw, len, vs, xfp + xfp_off[s] + (static_cast<int64_t>(j) * nhx + h) * kRB);
}

// y_c of column c of a big front after the forward pass: the sum of its
// partials.
template <bool C, class TX>
__device__ __forceinline__ TX forward_value(const TX* xfp_s, int nc, int c) {
  const int nhx = (nc + kXcRows - 1) / kXcRows;
  const int j = c / kRB;
  const int nv = (min(nc, (j + 1) * kRB) + kXcRows - 1) / kXcRows;
  const TX* pp = xfp_s + static_cast<int64_t>(j) * nhx * kRB + (c % kRB);
  TX v = TX(0);
  for (int h = 0; h < nv; ++h) v += ldx<C>(pp + static_cast<int64_t>(h) * kRB);
  return v;
}

// Row-block GEMV over a panel: returns sum_{c in [cb, ce)} M[r, c] y_c for the
// row r = lane (< rows) of the 64-row block at M, with y_c = stage(c) staged
// in LDS (smem: kSmem entries). Each lane owns a row, the four waves split
// the columns (16 loads in flight each), and the wave results are added in
// a fixed order.
template <class TF, class TX, class Stage>
__device__ __forceinline__ TX rowblock_gemv(const TF* __restrict__ M, int64_t ld, int rows, int cb, int ce,
                                            Stage stage, TX* smem) {
  TX* ys = smem;
  TX(*red)[kRB] = reinterpret_cast<TX(*)[kRB]>(smem + kChunk);
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
template <bool C, class TF, class TD, class TX>
__device__ __forceinline__ void fwd_u(const SolveArgs<TF, TD>& a, int s, int b, const int64_t* __restrict__ xfp_off,
                                      const TX* xfp, TX* uvec, TX* smem) {
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int64_t ld = panel_ld(nc, nr);
  const int rb0 = b * kRB;
  const int rows = min(kRB, nr - rb0);
  const TF* M = a.panel + a.panel_off[s] + nc + rb0;
  const TX* xs = xfp + xfp_off[s];
  const TX sum =
      rowblock_gemv<TF, TX>(M, ld, rows, 0, nc, [=](int c) { return forward_value<C>(xs, nc, c); }, smem);
  const int lane = threadIdx.x;
  if (lane < rows) {
    TX* u = uvec + a.rows_off[s] + rb0 + lane;
    stx<C>(u, ldx<C>(u) - sum);
  }
}

// Backward, rows below the columns of a front (R rows, all final): partial
// sums part[(j * nch + c) * 64 + t] = sum over the chunk's rows r of
// L[nc + r, 64 j + t] x[R_r].
template <bool C, class TF, class TD, class TX>
__device__ __forceinline__ void bwd_rpart(const SolveArgs<TF, TD>& a, int s, int j, int ch,
                                          const int64_t* __restrict__ part_off, const TX* x, TX* part, TX* xs) {
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int64_t ld = panel_ld(nc, nr);
  const int c0 = j * kRB;
  const int bw = min(kRB, nc - c0);
  const int nch = (nr + kRpartRows - 1) / kRpartRows;
  const int32_t* R = a.rows + a.rows_off[s];
  const int q0 = ch * kRpartRows, len = min(kRpartRows, nr - q0);
  for (int q = threadIdx.x; q < len; q += blockDim.x) xs[q] = ldx<C>(x + R[q0 + q]);
  __syncthreads();
  const TF* base = a.panel + a.panel_off[s] + nc + q0 + c0 * ld;
  column_dots<C, TF, TX>(base, ld, bw, len, xs,
                         part + part_off[s] + static_cast<int64_t>(j) * nch * kRB + static_cast<int64_t>(ch) * kRB);
}

// Backward, 64-row block i of the columns of front s:
// x_r = sum_{c >= 64 i} W[r, c] z_c, z = D^{-1} y - (partial sums of L_21^T x_R).
template <bool C, class TF, class TD, class TX>
__device__ __forceinline__ void bwd_xc(const SolveArgs<TF, TD>& a, int s, int i, const int64_t* __restrict__ part_off,
                                       const TX* part, const int64_t* __restrict__ xfp_off, const TX* xfp, TX* x,
                                       TX* smem) {
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
        const TX z = static_cast<TX>(Dinv[c]) * forward_value<false>(xs, nc, c);
        const TX* pp = ps + static_cast<int64_t>(c / kRB) * nch * kRB + (c % kRB);
        TX rs = TX(0);
        for (int h = 0; h < nch; ++h) rs += ldx<C>(pp + static_cast<int64_t>(h) * kRB);
        return z - rs;
      },
      smem);
  const int lane = threadIdx.x;
  if (lane < rows) stx<C>(x + f + r0 + lane, sum);
}

// ---------------------------------------------------------------- level schedule
template <class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_fwd_small(SolveArgs<TF, TD> a, const int32_t* __restrict__ list, TX* x,
                                                  TX* uvec) {
  __shared__ TX v[kSmallFront];
  small_fwd<false>(a, list[blockIdx.x], x, uvec, v, NoWait{});
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_bwd_small(SolveArgs<TF, TD> a, const int32_t* __restrict__ list, TX* x) {
  __shared__ TX w[kSmallFront];
  small_bwd<false>(a, list[blockIdx.x], x, w, NoWait{});
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_fwd_gather(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                                   const int32_t* __restrict__ cptr, const int32_t* __restrict__ cl,
                                                   TX* x, TX* uvec) {
  __shared__ TX vsh[kRB];
  gather_rows<false>(a, tasks[2 * blockIdx.x], tasks[2 * blockIdx.x + 1], cl, cptr[blockIdx.x], cptr[blockIdx.x + 1],
                     x, uvec, vsh);
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_fwd_xc(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                                const int64_t* __restrict__ xfp_off, const TX* x, TX* xfp) {
  __shared__ TX vs[kXcRows];
  fwd_xc<false>(a, tasks[3 * blockIdx.x], tasks[3 * blockIdx.x + 1], tasks[3 * blockIdx.x + 2], xfp_off, x, xfp, vs);
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_fwd_u(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                               const int64_t* __restrict__ xfp_off, const TX* xfp, TX* uvec) {
  __shared__ TX smem[kSmem];
  fwd_u<false>(a, tasks[2 * blockIdx.x], tasks[2 * blockIdx.x + 1], xfp_off, xfp, uvec, smem);
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_bwd_rpart(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                                   const int64_t* __restrict__ part_off, const TX* x, TX* part) {
  __shared__ TX xs[kRpartRows];
  bwd_rpart<false>(a, tasks[3 * blockIdx.x], tasks[3 * blockIdx.x + 1], tasks[3 * blockIdx.x + 2], part_off, x, part,
                   xs);
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_bwd_xc(SolveArgs<TF, TD> a, const int32_t* __restrict__ tasks,
                                                const int64_t* __restrict__ part_off, const TX* part,
                                                const int64_t* __restrict__ xfp_off, const TX* xfp, TX* x) {
  __shared__ TX smem[kSmem];
  bwd_xc<false>(a, tasks[2 * blockIdx.x], tasks[2 * blockIdx.x + 1], part_off, part, xfp_off, xfp, x, smem);
}

// ---------------------------------------------------------------- dag schedule
// Completion counters of front s (counters[k * ns + s]):
enum : int {
  kCDone = 0,  // children whose forward update vector is complete
  kGDone,      // forward gather tasks done
  kXDone,      // forward column-block tasks done
  kUDone,      // forward R-block tasks done
  kRDone,      // backward R-partial tasks done
  kBDone,      // backward column tasks done (small fronts: 1)
};
static_assert(kBDone + 1 == kNumSolveCounters, "counter layout");

// Waits until *c >= target. The wait is bounded: on expiry it records bit 1
// in *err and returns (the result is then wrong and the caller reports it).
__device__ __forceinline__ void spin_until(const int* c, int target, int* err) {
  int it = 0;
  while (ld_dev(c) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++it > kSpinLimit) {
      atomicOr(err, 2);
      break;
    }
  }
}
struct BlockWait {
  const int* c;
  int target;
  int* err;
  __device__ void operator()() const {
    if (threadIdx.x == 0) spin_until(c, target, err);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    __syncthreads();
  }
};
struct WaveWait {
  const int* c;
  int target;
  int* err;
  __device__ void operator()() const {
    if ((threadIdx.x & 63) == 0) spin_until(c, target, err);
    wave_sync();
  }
};
__device__ __forceinline__ int add_dev(int* c, int v) {
  return __hip_atomic_fetch_add(c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Called by every thread of the block after its last store of a task: once
// all those stores are complete, thread 0 counts the task and returns the
// counter's previous value (other threads: -1).
__device__ __forceinline__ int block_publish(int* c) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  return threadIdx.x == 0 ? add_dev(c, 1) : -1;
}
// The same for a task of one wavefront.
__device__ __forceinline__ void wave_publish(int* c) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync();
  if ((threadIdx.x & 63) == 0) add_dev(c, 1);
}
// The block's next task. The list is dealt round-robin onto kDispensers
// counters, each on its own 128-byte line (one counter serialises every
// workgroup of the device on one address); block b draws from counter
// b % kDispensers, so each counter hands out its share of the list in
// list order.
__device__ __forceinline__ int next_task(int* dispensers, int* tsh) {
  const int d = blockIdx.x % kDispensers;
  __syncthreads();
  if (threadIdx.x == 0) *tsh = add_dev(dispensers + kDispenserStride * d, 1) * kDispensers + d;
  __syncthreads();
  return *tsh;
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_fwd_dag(SolveArgs<TF, TD> a, DagArgs d, TX* x, TX* uvec, TX* xfp, int* err) {
  __shared__ TX smem[kSmem];
  __shared__ int tsh;
  const int w = threadIdx.x >> 6;
  int* cnt = d.counters;
  const int ns = d.ns;
  for (;;) {
    const int t = next_task(cnt + dispenser_base(ns), &tsh);
    if (t >= d.nfwd) break;
    const Task4 tk = d.fwd[t];
    if (tk.a == kDagSmall) {
      if (w < tk.c) {
        const int s = d.small_list[tk.b + w];
        const int* ms = d.meta + kMetaStride * s;
        small_fwd<true>(a, s, x, uvec, smem + kRB * w, WaveWait{cnt + kCDone * ns + s, ms[kMetaNChild], err});
        if (ms[kMetaParent] >= 0) wave_publish(cnt + kCDone * ns + ms[kMetaParent]);
      }
    } else if (tk.a == kDagGather) {
      if (w < tk.c) {
        const int g = tk.b + w;
        const int s = d.gather_task[2 * g], i = d.gather_task[2 * g + 1];
        WaveWait{cnt + kCDone * ns + s, d.meta[kMetaStride * s + kMetaNChild], err}();
        gather_rows<true>(a, s, i, d.gather_child, d.gather_child_ptr[g], d.gather_child_ptr[g + 1], x, uvec,
                          smem + kRB * w);
        wave_publish(cnt + kGDone * ns + s);
      }
    } else if (tk.a == kDagXc) {
      const int* q = d.xc_task + 3 * tk.b;
      const int s = q[0];
      BlockWait{cnt + kGDone * ns + s, d.meta[kMetaStride * s + kMetaNGather], err}();
      fwd_xc<true>(a, s, q[1], q[2], d.xfp_off, x, xfp, smem);
      block_publish(cnt + kXDone * ns + s);
    } else {  // kDagU
      const int* q = d.u_task + 2 * tk.b;
      const int s = q[0];
      const int* ms = d.meta + kMetaStride * s;
      BlockWait{cnt + kXDone * ns + s, ms[kMetaNXc], err}();
      fwd_u<true>(a, s, q[1], d.xfp_off, xfp, uvec, smem);
      // the front's last R block completes its update vector
      const int prev = block_publish(cnt + kUDone * ns + s);
      if (prev == ms[kMetaNU] - 1 && ms[kMetaParent] >= 0) add_dev(cnt + kCDone * ns + ms[kMetaParent], 1);
    }
  }
}

template <class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_bwd_dag(SolveArgs<TF, TD> a, DagArgs d, TX* x, TX* part, const TX* xfp,
                                                 int* err) {
  __shared__ TX smem[kSmem];
  __shared__ int tsh;
  const int w = threadIdx.x >> 6;
  int* cnt = d.counters;
  const int ns = d.ns;
  for (;;) {
    const int t = next_task(cnt + dispenser_base(ns) + kDispensers * kDispenserStride, &tsh);
    if (t >= d.nbwd) break;
    const Task4 tk = d.bwd[t];
    if (tk.a == kDagSmall) {
      if (w < tk.c) {
        const int s = d.small_list[tk.b + w];
        const int p = d.meta[kMetaStride * s + kMetaParent];
        // a root waits for nothing: its counter target is 0
        const int* pc = cnt + kBDone * ns + (p >= 0 ? p : s);
        const int target = p >= 0 ? d.meta[kMetaStride * p + kMetaNBFinal] : 0;
        small_bwd<true>(a, s, x, smem + kRB * w, WaveWait{pc, target, err});
        wave_publish(cnt + kBDone * ns + s);
      }
    } else if (tk.a == kDagRpart) {
      const int* q = d.rpart_task + 3 * tk.b;
      const int s = q[0];
      const int p = d.meta[kMetaStride * s + kMetaParent];
      if (p >= 0) BlockWait{cnt + kBDone * ns + p, d.meta[kMetaStride * p + kMetaNBFinal], err}();
      bwd_rpart<true>(a, s, q[1], q[2], d.part_off, x, part, smem);
      block_publish(cnt + kRDone * ns + s);
    } else {  // kDagBxc
      const int* q = d.bwd_task + 2 * tk.b;
      const int s = q[0];
      BlockWait{cnt + kRDone * ns + s, d.meta[kMetaStride * s + kMetaNRpart], err}();
      bwd_xc<true>(a, s, q[1], d.part_off, part, d.xfp_off, xfp, x, smem);
      block_publish(cnt + kBDone * ns + s);
    }
  }
}

__global__ void k_zero(int n, int* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Workgroups of a kernel resident on the whole device at once.
template <class K>
int resident_blocks(K kernel, int threads) {
  int dev = 0, cus = 0, per_cu = 0;
  MPIR_HIP(hipGetDevice(&dev));
  MPIR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  MPIR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  return std::max(1, per_cu) * std::max(1, cus);
}

}  // namespace

template <class TF, class TD, class TX>
void ldl_solve(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st) {
  TX* xfp = reinterpret_cast<TX*>(dp.xfp);
  TX* part = reinterpret_cast<TX*>(dp.part);
  if (dp.schedule == kScheduleDag) {
    // Every task waits only on tasks before it in the list. The smallest
    // unfinished task is therefore either running or next on its
    // dispenser, whose earlier tasks are all done, so a block of that
    // dispenser takes it: blocks 0 .. kDispensers - 1 are dispatched first
    // and serve every dispenser until its share of the list runs out.
    static const int grid_f = resident_blocks(k_fwd_dag<TF, TD, TX>, 256);
    static const int grid_b = resident_blocks(k_bwd_dag<TF, TD, TX>, 256);
    const DagArgs& d = dp.dag;
    const int nz = dag_counter_count(d.ns);
    hipLaunchKernelGGL(k_zero, dim3((nz + 255) / 256), dim3(256), 0, st, nz, d.counters);
    if (d.nfwd > 0)
      hipLaunchKernelGGL((k_fwd_dag<TF, TD, TX>), dim3(std::max(kDispensers, std::min(d.nfwd, grid_f))), dim3(256), 0,
                         st, a, d, x, uvec, xfp, dp.err);
    if (d.nbwd > 0)
      hipLaunchKernelGGL((k_bwd_dag<TF, TD, TX>), dim3(std::max(kDispensers, std::min(d.nbwd, grid_b))), dim3(256), 0,
                         st, a, d, x, part, static_cast<const TX*>(xfp), dp.err);
    MPIR_HIP(hipGetLastError());
    return;
  }
  // forward, leaves to root; D^{-1} fused
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_fwd_small<TF, TD, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x, uvec);
    const int g0 = dp.h_gather_ptr[l], g1 = dp.h_gather_ptr[l + 1];
    if (g1 > g0)
      hipLaunchKernelGGL((k_fwd_gather<TF, TD, TX>), dim3(g1 - g0), dim3(64), 0, st, a, dp.gather_task + 2 * g0,
                         dp.gather_child_ptr + g0, dp.gather_child, x, uvec);
    const int c0 = dp.h_xc_ptr[l], c1 = dp.h_xc_ptr[l + 1];
    if (c1 > c0)
      hipLaunchKernelGGL((k_fwd_xc<TF, TD, TX>), dim3(c1 - c0), dim3(256), 0, st, a, dp.xc_task + 3 * c0, dp.xfp_off,
                         x, xfp);
    const int u0 = dp.h_u_ptr[l], u1 = dp.h_u_ptr[l + 1];
    if (u1 > u0)
      hipLaunchKernelGGL((k_fwd_u<TF, TD, TX>), dim3(u1 - u0), dim3(256), 0, st, a, dp.u_task + 2 * u0, dp.xfp_off,
                         xfp, uvec);
  }
  // backward, root to leaves
  for (int l = dp.nlevels - 1; l >= 0; --l) {
    const int p0 = dp.h_rpart_ptr[l], p1 = dp.h_rpart_ptr[l + 1];
    if (p1 > p0)
      hipLaunchKernelGGL((k_bwd_rpart<TF, TD, TX>), dim3(p1 - p0), dim3(256), 0, st, a, dp.rpart_task + 3 * p0,
                         dp.part_off, x, part);
    const int t0 = dp.h_bwd_ptr[l], t1 = dp.h_bwd_ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL((k_bwd_xc<TF, TD, TX>), dim3(t1 - t0), dim3(256), 0, st, a, dp.bwd_task + 2 * t0,
                         dp.part_off, part, dp.xfp_off, xfp, x);
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_bwd_small<TF, TD, TX>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0, x);
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
