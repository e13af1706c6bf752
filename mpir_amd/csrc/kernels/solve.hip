// Supernodal triangular solves with the multifrontal factor:
//   forward  (L + I) y = b,  y := D^{-1} y,  backward (L + I)^T x = y.
// This is QDLDL_solve (include/MPIR/qdldl.hpp:417-432 of yex33/MPIR: Lsolve
// :362-372, the Dinv step :430, Ltsolve :389-400) with the factor held in
// factor precision TF and every product and sum done in the precision of x,
// TX, as QDLDL_Lsolve/QDLDL_Ltsolve do with XPREC(Lx[j]).
//
// Small fronts (m <= 64, one wavefront each): gather b and the children's
// update vectors (extend-add), solve the unit-lower block, leave
// u_s = v_R - L_21 y for the parent; backward, gather x on R_s (all
// ancestors, already final) and solve the block transposed.
//
// Big fronts hold W = L_11^{-T} (unit upper triangular) below their panel,
// so their solves are matrix-vector products without a dependency chain:
//   forward   y = W^T v,  u = v_R - L_21 y,  and D^{-1} y is kept
//   backward  z = D^{-1} y - L_21^T x_R,  x = W z
// Each product is cut into 16K-entry tiles (kXcRows rows x 64 columns for
// W^T v and L_21^T x_R, 64 rows x kGemvCols columns for L_21 y and W z)
// read with 16-byte loads; a tile is loaded before its task waits for its
// inputs. Tiles leave partial sums; the last tile of a block to finish
// (counted on a per-block counter) adds the block's partials in a fixed
// order, so repeated solves agree bit for bit. The factor is stored in TF
// (double, float, half or bfloat16), D^{-1} in TD (TF, or float for the
// 16-bit factors); products and sums run in TX.
//
// Two schedules run the same task bodies, so their results agree bit for
// bit:
//   levels: one launch per task kind and level of the supernodal tree;
//   dag:    one persistent launch per direction. Workgroups take tasks in a
//           topological order from a set of dispensers and wait (bounded)
//           on per-front completion counters, so a front starts as soon as
//           its own children (forward) or its parent (backward) are done.
// Values passed between workgroups use device-coherent stores and loads,
// and a task is counted only after its stores have completed.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "../common.hpp"
#include "kernels.hpp"
#include "sync.hpp"

namespace mpir {

namespace {

constexpr int kRB = kSolveRB;
// LDS of a 256-thread task, in entries of the solve precision: the staged
// vector of a tile (up to 512 rows for the column products of 16-bit
// factors, kGemvCols entries for the row products), then the wave
// reduction of the row products (4 x 64)
constexpr int kVsLen = 512;
static_assert(kXcRows <= kVsLen && kGemvCols <= kVsLen, "staging buffer");

// Per-front completion counters (counters[k * ns + s]).
enum : int {
  kCDone = 0,  // children whose update vector is complete
  kXDone,      // forward column-block tasks (partial sums of y) done
  kUDone,      // forward R-block tasks done
  kRDone,      // backward R-chunk tasks done
  kBDone,      // backward column blocks with x final (small fronts: 1)
};
static_assert(kBDone + 1 == kNumSolveCounters, "counter layout");

// 16 bytes of factor entries; the address needs only the element's alignment
template <class TF>
struct Pack {
  TF e[16 / sizeof(TF)];
};
template <class TF>
__device__ __forceinline__ Pack<TF> load_pack(const TF* p) {
  Pack<TF> r;
  __builtin_memcpy(&r, p, sizeof r);
  return r;
}

// sum_{k < n} p[k * stride], added in order; the loads go out 8 at a time
template <class TX>
__device__ __forceinline__ TX sum_strided(const TX* p, int n, int stride) {
  TX s = TX(0);
  for (int k0 = 0; k0 < n; k0 += 8) {
    TX t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = k0 + k < n ? ld_dev(p + static_cast<int64_t>(k0 + k) * stride) : TX(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) s += t[k];
  }
  return s;
}

// s[r] = sum_{k < n} p[r * rs + k * stride] for NR right-hand sides, each
// added in order as sum_strided does; the loads of all right-hand sides go
// out together, 8 at a time (k outer, r inner)
template <int NR, class TX>
__device__ __forceinline__ void sum_strided_multi(const TX* p, int n, int stride, int64_t rs, TX (&s)[NR]) {
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = TX(0);
  if constexpr (NR == 1) {
    s[0] = sum_strided(p, n, stride);
  } else {
    const int total = n * NR;
    for (int i0 = 0; i0 < total; i0 += 8) {
      TX t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u, k = i / NR, r = i % NR;
        t[u] = i < total ? ld_dev(p + r * rs + static_cast<int64_t>(k) * stride) : TX(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (i0 + u < total) s[(i0 + u) % NR] += t[u];
    }
    // sum_strided adds the zeros of its last partial batch, which turns a
    // -0 sum into +0: the same here, so the bits agree
    if (n % 8 != 0)
#pragma unroll
      for (int r = 0; r < NR; ++r) s[r] += TX(0);
  }
}

// y of column c of a big front: its partial sums over the row chunks
template <class TX>
__device__ __forceinline__ TX forward_value(const TX* yp, int nc, int c) {
  const int nhx = (nc + kXcRows - 1) / kXcRows;
  const int j = c / kSolveRB;
  const int nv = (min(nc, (j + 1) * kSolveRB) + kXcRows - 1) / kXcRows;
  return sum_strided(yp + static_cast<int64_t>(j) * nhx * kSolveRB + (c % kSolveRB), nv, kSolveRB);
}
// the same for NR right-hand sides at stride rs
template <int NR, class TX>
__device__ __forceinline__ void forward_values(const TX* yp, int64_t rs, int nc, int c, TX (&y)[NR]) {
  const int nhx = (nc + kXcRows - 1) / kXcRows;
  const int j = c / kSolveRB;
  const int nv = (min(nc, (j + 1) * kSolveRB) + kXcRows - 1) / kXcRows;
  sum_strided_multi<NR>(yp + static_cast<int64_t>(j) * nhx * kSolveRB + (c % kSolveRB), nv, kSolveRB, rs, y);
}

// The children's contributions to one front row (entries [p0, p1) of the
// gather table, in child order): the child's update-vector entry less the
// partial sums of L_21 y when the child is big. The first two entries are
// loaded by pre(), before the task waits; value() loads the values.
struct RowGather {
  int p0 = 0, p1 = 0;
  Task4 e0{0, 0, 0, 0}, e1{0, 0, 0, 0};
  __device__ __forceinline__ void pre(const SolveDev& d, int frow) {
    p0 = d.grow_ptr[frow];
    p1 = d.grow_ptr[frow + 1];
    e0 = d.gent[p0];
    e1 = d.gent[p0 + 1];
  }
  // uvec, up: the update vectors and L_21 y partial sums of one right-hand side
  template <class TX>
  __device__ __forceinline__ TX value(const SolveDev& d, const TX* uvec, const TX* up, TX v) const {
    for (int p = p0; p < p1; ++p) {
      const Task4 e = p == p0 ? e0 : p == p0 + 1 ? e1 : d.gent[p];
      v += ld_dev(uvec + e.a) - sum_strided(up + e.b, e.c, kSolveRB);
    }
    return v;
  }
  // the same for NR right-hand sides (stride d.su, d.sup), v[r] in and out:
  // each entry's loads go out for every right-hand side at once, and each
  // right-hand side adds them in the order value() does
  template <int NR, class TX>
  __device__ __forceinline__ void values(const SolveDev& d, const TX* uvec, const TX* up, TX (&v)[NR]) const {
    if constexpr (NR == 1) {
      v[0] = value(d, uvec, up, v[0]);
    } else {
      for (int p = p0; p < p1; ++p) {
        const Task4 e = p == p0 ? e0 : p == p0 + 1 ? e1 : d.gent[p];
        TX u[NR], ps[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) u[r] = ld_dev(uvec + r * d.su + e.a);
        sum_strided_multi<NR>(up + e.b, e.c, kSolveRB, d.sup, ps);
#pragma unroll
        for (int r = 0; r < NR; ++r) v[r] += u[r] - ps[r];
      }
    }
  }
};

// ---------------------------------------------------------------- small fronts
// One wavefront per front (m <= 64), lane = row; v: 64 entries of LDS. L is
// read kLC columns (rows) at a time, all loads of a group in flight at once.
// wait() runs after the front's first loads and before anything another
// task writes is read.
constexpr int kLC = 16;

template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void small_fwd(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, TX* x, TX* uvec,
                                          Wait wait) {
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
  RowGather G;
  if (lane < m) G.pre(d, d.frow_off[s] + lane);
  TX b[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) b[r] = x[r * d.sx + f + min(lane, nc - 1)];
  const TD dinv = a.Dinv[f + min(lane, nc - 1)];
  wait();
  const TX* up = static_cast<const TX*>(d.upart);
  TX vl[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) vl[r] = lane < nc ? b[r] : TX(0);
  G.values<NR>(d, uvec, up, vl);
  for (int j0 = 0; j0 < nc; j0 += kLC) {
    if (j0 > 0) load(j0);
#pragma unroll
    for (int u = 0; u < kLC; ++u) {
      const int j = j0 + u;
      if (j < nc) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const TX yj = lane_value(vl[r], j);
          if (lane > j) vl[r] -= static_cast<TX>(lr[u]) * yj;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (lane < nc) x[r * d.sx + f + lane] = vl[r] * static_cast<TX>(dinv);
    else if (lane < m) st_dev(uvec + r * d.su + a.rows_off[s] + (lane - nc), vl[r]);
  }
}

// w: NR x 64 entries of LDS
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void small_bwd(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, TX* x, TX* w,
                                          Wait wait) {
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
  TX z[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) z[r] = x[r * d.sx + f + cl];
  wait();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (lane < nc) w[r * kRB + lane] = z[r];
    else if (lane < m) w[r * kRB + lane] = ld_dev(x + r * d.sx + row);
  }
  wave_sync();
  // lane j < nc: z_j = w_j - sum over rows r >= nc of L[r, j] w_r, r ascending
  for (int r0 = nc; r0 < m; r0 += kLC) {
    if (r0 > nc) load_up(r0);
#pragma unroll
    for (int u = 0; u < kLC; ++u)
      if (r0 + u < m)
#pragma unroll
        for (int r = 0; r < NR; ++r) z[r] -= static_cast<TX>(lc[u]) * w[r * kRB + r0 + u];
  }
  // unit upper solve with L_11^T, bottom to top
  for (int t1 = nc; t1 > 0; t1 -= kLC) {
    load_down(t1);
#pragma unroll
    for (int u = 0; u < kLC; ++u) {
      const int t = t1 - 1 - u;
      if (t >= 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const TX xt = lane_value(z[r], t);
          if (lane < t) z[r] -= static_cast<TX>(lc[u]) * xt;
        }
      }
    }
  }
  if (lane < nc)
#pragma unroll
    for (int r = 0; r < NR; ++r) st_dev(x + r * d.sx + f + lane, z[r]);
}

// ---------------------------------------------------------------- big-front tiles
// Column products: up to kXcRows rows x 64 columns, entry (q, t) at
// base[q + t * ld]; out[t] = sum_q M[q, t] vs[q]. A lane loads N = 16 /
// sizeof(TF) consecutive rows, a wave load covers 64 N rows of one column,
// and wave w takes columns 16 w .. 16 w + 15. Loads go in batches of at
// most 16 per lane; batch 0 is issued before the task's wait.
template <class TF>
struct ColTile {
  static constexpr int N = 16 / sizeof(TF);
  static constexpr int RPL = 64 * N;                     // rows per wave load
  static constexpr int NLC = (kXcRows + RPL - 1) / RPL;  // loads per column
  static constexpr int NL = 16 * NLC;                    // loads per lane
  static constexpr int B = NL < 16 ? NL : 16;
  static constexpr int NB = NL / B;
  static_assert(NL % B == 0 && NLC * RPL <= kVsLen, "tile shape");
  Pack<TF> buf[B];
  const TF* base;
  int64_t ld;
  int bw, len;
  __device__ __forceinline__ void load(int batch) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int k = batch * B + u;
      const int t = min(w * 16 + k / NLC, bw - 1);
      const int r = (k % NLC) * RPL + N * lane;
      // rows past the tile read the tile's first rows instead (masked)
      buf[u] = load_pack(base + (r < len ? r : 0) + t * ld);
    }
  }
};

// Sums of acc[k] over the 64 lanes of a wave, k < 16, by halving exchanges
// (fixed order: 17 exchanges instead of 16 x 6): after the step with offset
// o a lane keeps the half of its entries selected by its bit o. Returns the
// sum of entry c (set here) in the 4 lanes 4 c', 4 c' + 1, .. of each group.
template <class TX>
__device__ __forceinline__ TX wave_sum16(const TX (&acc)[16], int& c) {
  const int lane = threadIdx.x & 63;
  TX a8[8], a4[4], a2[2];
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
  for (int i = 0; i < 8; ++i) a8[i] = (b5 ? acc[i + 8] : acc[i]) + __shfl_xor(b5 ? acc[i] : acc[i + 8], 32);
#pragma unroll
  for (int i = 0; i < 4; ++i) a4[i] = (b4 ? a8[i + 4] : a8[i]) + __shfl_xor(b4 ? a8[i] : a8[i + 4], 16);
#pragma unroll
  for (int i = 0; i < 2; ++i) a2[i] = (b3 ? a4[i + 2] : a4[i]) + __shfl_xor(b3 ? a4[i] : a4[i + 2], 8);
  TX a1 = (b2 ? a2[1] : a2[0]) + __shfl_xor(b2 ? a2[0] : a2[1], 4);
  a1 += __shfl_xor(a1, 2);
  a1 += __shfl_xor(a1, 1);
  c = (b5 ? 8 : 0) + (b4 ? 4 : 0) + (b3 ? 2 : 0) + (b2 ? 1 : 0);
  return a1;
}

// The tile's column sums (vs: zeros from len to kVsLen) stored to out[t],
// t < bw. The 16 sums of each lane are reduced across the wave by
// wave_sum16.
template <class TF, class TX>
__device__ __forceinline__ void col_dots(ColTile<TF>& T, const TX* vs, TX* out) {
  using CT = ColTile<TF>;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  TX v[CT::NLC][CT::N];
  bool ok[CT::NLC][CT::N];
#pragma unroll
  for (int g = 0; g < CT::NLC; ++g)
#pragma unroll
    for (int e = 0; e < CT::N; ++e) {
      const int r = g * CT::RPL + CT::N * lane + e;
      v[g][e] = vs[r];
      ok[g][e] = r < T.len;
    }
  TX acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = TX(0);
#pragma unroll
  for (int b = 0; b < CT::NB; ++b) {
    if (b > 0) T.load(b);
#pragma unroll
    for (int u = 0; u < CT::B; ++u) {
      const int k = b * CT::B + u, c = k / CT::NLC, g = k % CT::NLC;
#pragma unroll
      for (int e = 0; e < CT::N; ++e) acc[c] += (ok[g][e] ? static_cast<TX>(T.buf[u].e[e]) : TX(0)) * v[g][e];
    }
  }
  int c = 0;
  const TX a1 = wave_sum16(acc, c);
  if ((lane & 3) == 0 && w * 16 + c < T.bw) st_dev(out + w * 16 + c, a1);
}

// Row products: 64 rows x up to kGemvCols columns, entry (r, c) at
// base[r + c * ld]; sum_c M[r, c] ys[c]. A wave load covers N columns of 64
// rows (lane l: rows N (l % LPC) .., column l / LPC of the group); wave w
// takes columns 64 w .. 64 w + 63. Loads go in batches of at most 16 per
// lane; batch 0 is issued before the task's wait.
template <class TF>
struct RowTile {
  static constexpr int N = 16 / sizeof(TF);
  static constexpr int LPC = 64 / N;  // lanes per column
  static constexpr int NL = 64 / N;   // loads per lane
  static constexpr int B = NL < 16 ? NL : 16;
  static constexpr int NB = NL / B;
  static_assert(NL % B == 0 && 4 * 64 == kGemvCols, "tile shape");
  Pack<TF> buf[B];
  const TF* base;
  int64_t ld;
  int rows, cc;
  __device__ __forceinline__ void load(int batch) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = N * (lane % LPC);
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int g = batch * B + u;
      const int c = min(w * 64 + g * N + lane / LPC, cc - 1);
      // rows past the tile read the tile's first rows instead (results unused)
      buf[u] = load_pack(base + (r < rows ? r : 0) + c * ld);
    }
  }
};

// The tile's row sums, one right-hand side's part: each wave's sums of its
// columns go to red[w][row] (4 x 64 entries of LDS, one such slot per
// right-hand side, so that the parts of several right-hand sides need one
// barrier between them and row_sum). ys: kGemvCols entries.
template <class TF, class TX>
__device__ __forceinline__ void row_gemv_part(RowTile<TF>& T, const TX* ys, TX (*red)[kRB]) {
  using RT = RowTile<TF>;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  TX acc[RT::N];
#pragma unroll
  for (int e = 0; e < RT::N; ++e) acc[e] = TX(0);
#pragma unroll
  for (int b = 0; b < RT::NB; ++b) {
    if (b > 0) T.load(b);
#pragma unroll
    for (int u = 0; u < RT::B; ++u) {
      const int c = w * 64 + (b * RT::B + u) * RT::N + lane / RT::LPC;
      const bool ok = c < T.cc;
      const TX y = ys[c];
#pragma unroll
      for (int e = 0; e < RT::N; ++e) acc[e] += (ok ? static_cast<TX>(T.buf[u].e[e]) : TX(0)) * y;
    }
  }
  // lanes l, l + LPC, ... hold the same rows
#pragma unroll
  for (int off = RT::LPC; off < 64; off <<= 1)
#pragma unroll
    for (int e = 0; e < RT::N; ++e) acc[e] += __shfl_xor(acc[e], off);
  if (lane < RT::LPC)
#pragma unroll
    for (int e = 0; e < RT::N; ++e) red[w][RT::N * lane + e] = acc[e];
}
// the row sums for threads 0..63 (other threads: 0), after a barrier that
// follows row_gemv_part
template <class TX>
__device__ __forceinline__ TX row_sum(const TX (*red)[kRB]) {
  const int r = threadIdx.x;
  return r < kRB ? (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]) : TX(0);
}

// ---------------------------------------------------------------- big-front tasks
// f(r) for the right-hand sides r < NR (one right-hand side: no loop, so the
// single case compiles exactly as written for it)
template <int NR, class F>
__device__ __forceinline__ void each_rhs(F&& f) {
  if constexpr (NR == 1) {
    f(0);
  } else {
#pragma unroll 1
    for (int r = 0; r < NR; ++r) f(r);
  }
}

struct FrontShape {
  int f, nc, nr, m;
  int64_t ld;
  template <class TF, class TD>
  __device__ __forceinline__ FrontShape(const SolveArgs<TF, TD>& a, int s) {
    f = a.first[s];
    nc = a.first[s + 1] - f;
    nr = a.nrow[s];
    m = nc + nr;
    ld = panel_ld(nc, nr);
  }
};

// Forward, rows [h kXcRows, ...) of column block j: gathers v on those rows
// (b plus the children's contributions) and leaves partial sums of W^T v,
// for NR right-hand sides (the tile is read once).
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void fx_task(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, int j, int h, const TX* x,
                                        const TX* uvec, TX* smem, Wait wait) {
  static_assert(kXcRows == 256, "one row per thread");
  static_assert(NR == 1 || ColTile<TF>::NB == 1, "several right-hand sides need the tile in registers");
  const FrontShape F(a, s);
  const int tid = threadIdx.x;
  const int c0 = j * kRB, bw = min(kRB, F.nc - c0);
  const int q0 = h * kXcRows, len = min(kXcRows, c0 + bw - q0);
  ColTile<TF> T;
  T.base = a.panel + a.panel_off[s] + F.m + q0 + c0 * F.ld;
  T.ld = F.ld;
  T.bw = bw;
  T.len = len;
  T.load(0);
  RowGather G;
  TX bq[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) bq[r] = TX(0);
  if (tid < len) {
    G.pre(d, d.frow_off[s] + q0 + tid);
#pragma unroll
    for (int r = 0; r < NR; ++r) bq[r] = x[r * d.sx + F.f + q0 + tid];
  }
  wait();
  const TX* up = static_cast<const TX*>(d.upart);
  if (tid < len) G.values<NR>(d, uvec, up, bq);
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    TX* vs = smem + r * kVsLen;
    vs[tid] = tid < len ? bq[r] : TX(0);
    for (int q = tid + blockDim.x; q < kVsLen; q += blockDim.x) vs[q] = TX(0);
  }
  __syncthreads();
  const int nhx = (F.nc + kXcRows - 1) / kXcRows;
  TX* yp = static_cast<TX*>(d.ypart) + d.ypart_off[s] + (static_cast<int64_t>(j) * nhx + h) * kRB;
  each_rhs<NR>([&](int r) { col_dots<TF, TX>(T, smem + r * kVsLen, yp + r * d.syp); });
  block_publish(d.counters + kXDone * d.ns + s);
}

// Forward, R block b, columns [h kGemvCols, ...): partial sums of L_21 y;
// the chunk-0 task also gathers and stores v_R, so that u = v_R - (partial
// sums) is formed by whoever gathers it. The front's last task counts the
// front as a finished child of its parent. NR right-hand sides.
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void fu_task(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, int b, int h, TX* uvec,
                                        TX* smem, Wait wait) {
  static_assert(NR == 1 || RowTile<TF>::NB == 1, "several right-hand sides need the tile in registers");
  const FrontShape F(a, s);
  const int tid = threadIdx.x;
  const int rb0 = b * kRB, rows = min(kRB, F.nr - rb0);
  const int cs = h * kGemvCols, cc = min(kGemvCols, F.nc - cs);
  RowTile<TF> T;
  T.base = a.panel + a.panel_off[s] + F.nc + rb0 + cs * F.ld;
  T.ld = F.ld;
  T.rows = rows;
  T.cc = cc;
  T.load(0);
  const bool gathers = h == 0 && tid < rows;
  RowGather G;
  if (gathers) G.pre(d, d.frow_off[s] + F.nc + rb0 + tid);
  wait();
  const TX* yp = static_cast<const TX*>(d.ypart) + d.ypart_off[s];
  TX* upart = static_cast<TX*>(d.upart);
  TX vr[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) vr[r] = TX(0);
  for (int c = tid; c < kGemvCols; c += blockDim.x) {
    TX yv[NR];
    if (c < cc) forward_values<NR>(yp, d.syp, F.nc, cs + c, yv);
#pragma unroll
    for (int r = 0; r < NR; ++r) smem[r * kVsLen + c] = c < cc ? yv[r] : TX(0);
  }
  if (gathers) G.values<NR>(d, static_cast<const TX*>(uvec), upart, vr);
  __syncthreads();
  const int nh = (F.nc + kGemvCols - 1) / kGemvCols;
  TX* out = upart + d.upart_off[s] + static_cast<int64_t>(b) * nh * kRB + h * kRB;
  TX (*red)[4][kRB] = reinterpret_cast<TX(*)[4][kRB]>(smem + NR * kVsLen);
  each_rhs<NR>([&](int r) { row_gemv_part<TF, TX>(T, smem + r * kVsLen, red[r]); });
  __syncthreads();
  each_rhs<NR>([&](int r) {
    const TX sum = row_sum<TX>(red[r]);
    if (tid < rows) st_dev(out + r * d.sup + tid, sum);
  });
  if (gathers)
#pragma unroll
    for (int r = 0; r < NR; ++r) st_dev(uvec + r * d.su + a.rows_off[s] + rb0 + tid, vr[r]);
  const int prev = block_publish(d.counters + kUDone * d.ns + s);
  const int* ms = d.meta + kMetaStride * s;
  if (prev == ms[kMetaNFu] - 1 && ms[kMetaParent] >= 0) add_dev(d.counters + kCDone * d.ns + ms[kMetaParent], 1);
}

// Backward, R rows [h kXcRows, ...) of column block j: partial sums of
// L_21^T x_R for NR right-hand sides.
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void br_task(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, int j, int h,
                                        const TX* x, TX* smem, Wait wait) {
  static_assert(NR == 1 || ColTile<TF>::NB == 1, "several right-hand sides need the tile in registers");
  const FrontShape F(a, s);
  const int tid = threadIdx.x;
  const int c0 = j * kRB, bw = min(kRB, F.nc - c0);
  const int q0 = h * kXcRows, len = min(kXcRows, F.nr - q0);
  ColTile<TF> T;
  T.base = a.panel + a.panel_off[s] + F.nc + q0 + c0 * F.ld;
  T.ld = F.ld;
  T.bw = bw;
  T.len = len;
  T.load(0);
  wait();
  const int32_t* R = a.rows + a.rows_off[s] + q0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    TX* vs = smem + r * kVsLen;
    for (int q = tid; q < kVsLen; q += blockDim.x) vs[q] = q < len ? ld_dev(x + r * d.sx + R[q]) : TX(0);
  }
  __syncthreads();
  const int nhr = (F.nr + kXcRows - 1) / kXcRows;
  TX* rp = static_cast<TX*>(d.rpart) + d.rpart_off[s] + (static_cast<int64_t>(j) * nhr + h) * kRB;
  each_rhs<NR>([&](int r) { col_dots<TF, TX>(T, smem + r * kVsLen, rp + r * d.srp); });
  block_publish(d.counters + kRDone * d.ns + s);
}

// Backward, row block i, columns [max(h kGemvCols, 64 i), ...): partial sums
// of W z with z = D^{-1} y - L_21^T x_R; the block's last task stores x.
// NR right-hand sides.
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void bx_task(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, int i, int h, TX* x,
                                        TX* smem, int* flag, Wait wait) {
  static_assert(NR == 1 || RowTile<TF>::NB == 1, "several right-hand sides need the tile in registers");
  const FrontShape F(a, s);
  const int tid = threadIdx.x;
  const int r0 = i * kRB, rows = min(kRB, F.nc - r0);
  const int cs = max(h * kGemvCols, r0), cc = min(F.nc, (h + 1) * kGemvCols) - cs;
  RowTile<TF> T;
  T.base = a.panel + a.panel_off[s] + F.m + r0 + cs * F.ld;
  T.ld = F.ld;
  T.rows = rows;
  T.cc = cc;
  T.load(0);
  wait();
  const int nhr = (F.nr + kXcRows - 1) / kXcRows;
  const TX* rp = static_cast<const TX*>(d.rpart) + (nhr > 0 ? d.rpart_off[s] : 0);
  const TX* yp = static_cast<const TX*>(d.ypart) + d.ypart_off[s];
  for (int c = tid; c < kGemvCols; c += blockDim.x) {
    TX rs[NR], yv[NR];
    const int col = cs + c;
    if (c < cc) {
      sum_strided_multi<NR>(rp + static_cast<int64_t>(col / kRB) * nhr * kRB + (col % kRB), nhr, kRB, d.srp, rs);
      forward_values<NR>(yp, d.syp, F.nc, col, yv);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
      smem[r * kVsLen + c] = c < cc ? static_cast<TX>(a.Dinv[F.f + col]) * yv[r] - rs[r] : TX(0);
  }
  __syncthreads();
  const int nh = (F.nc + kGemvCols - 1) / kGemvCols;
  const int h0 = r0 / kGemvCols;
  TX* xp = static_cast<TX*>(d.xpart) + d.xpart_off[s] + static_cast<int64_t>(i) * nh * kRB;
  TX (*red)[4][kRB] = reinterpret_cast<TX(*)[4][kRB]>(smem + NR * kVsLen);
  each_rhs<NR>([&](int r) { row_gemv_part<TF, TX>(T, smem + r * kVsLen, red[r]); });
  __syncthreads();
  each_rhs<NR>([&](int r) {
    const TX sum = row_sum<TX>(red[r]);
    if (tid < rows) st_dev(xp + r * d.sxp + h * kRB + tid, sum);
  });
  if (publish_last(d.counters + d.cnt_bx + d.cb_base[s] + i, nh - h0, flag)) {
    if (tid < rows)
#pragma unroll
      for (int r = 0; r < NR; ++r)
        st_dev(x + r * d.sx + F.f + r0 + tid, sum_strided(xp + r * d.sxp + h0 * kRB + tid, nh - h0, kRB));
    block_publish(d.counters + kBDone * d.ns + s);
  }
}

// ---------------------------------------------------------------- narrow fronts
// A big front with nc <= 64 columns and at most a few kNarrowChunk-row
// chunks of R (the plan's narrow fronts). Its forward tasks (one per chunk)
// each form y = W^T v1 themselves (W is at most 64 x 64) and leave the
// final u = v_R - L_21 y of their rows, so no task waits on another task of
// its own front; its backward task forms x1 = W (D^{-1} y - L_21^T x_R)
// alone. Every reduction runs in a fixed order.
//
// Tile entries outside the front (columns past nc, rows past the chunk)
// read clamped or neighbouring addresses inside the panel pool, whose
// slack past the last front is zero: they are finite, and each meets a
// zero entry of v1, y, x_R or z or lands in a sum that is not stored, so
// no entry is masked (a factor with a non-finite entry gives a non-finite
// result either way).
//
// W (nc x nc, at most 64 x 64) in 16-byte packs: lane l of wave w holds
// N = 16 / sizeof(TF) consecutive rows N (l % 16) + 16 N p of the columns
// 16 k + 4 w + l / 16 (k < 4, p < P = 4 / N rounded up). Columns past the
// front read its last column, rows past it the next entries in memory.
template <class TF>
struct WTile {
  static constexpr int N = 16 / sizeof(TF);
  static constexpr int P = (64 + 16 * N - 1) / (16 * N);
  Pack<TF> e[4][P];
  __device__ __forceinline__ static int row(int p) { return N * (threadIdx.x & 15) + 16 * N * p; }
  __device__ __forceinline__ static int col(int k) { return 16 * k + 4 * (threadIdx.x >> 6) + ((threadIdx.x & 63) >> 4); }
  __device__ __forceinline__ void load(const TF* base, int64_t ld, int nc) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int r = row(p);
        e[k][p] = load_pack(base + (r < nc ? r : 0) + static_cast<int64_t>(min(col(k), nc - 1)) * ld);
      }
  }
};

// L_21 rows of a narrow front, kNarrowChunk rows x nc <= 64 columns, entry
// (r, c) at base[r + c * ld]: wave w takes rows 64 w .. 64 w + 63; lane l
// loads N = 16 / sizeof(TF) consecutive rows 64 w + N (l % LPC) .. of the
// columns N g + l / LPC (LPC = 64 / N lanes per column, load g < 64 / N).
// Loads go in batches of at most MB per lane (batch b: columns from
// N B b on).
template <class TF, int MB = 16>
struct NarrowTile {
  static constexpr int N = 16 / sizeof(TF);
  static constexpr int LPC = 64 / N;  // lanes per column
  static constexpr int NL = 64 / N;   // loads per lane
  static constexpr int B = NL < MB ? NL : MB;
  static constexpr int NB = NL / B;
  static_assert(NL % B == 0 && 4 * 64 == kNarrowChunk, "tile shape");
  Pack<TF> buf[B];
  const TF* base;
  int64_t ld;
  int rows, nc;
  __device__ __forceinline__ int row0() const { return 64 * (threadIdx.x >> 6) + N * ((threadIdx.x & 63) % LPC); }
  __device__ __forceinline__ int col(int g) const { return N * g + (threadIdx.x & 63) / LPC; }
  __device__ __forceinline__ void load(int batch) {
    const int r = row0();
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int c = min(col(batch * B + u), nc - 1);
      // rows past the chunk read its first rows instead (their sums are
      // not stored, or they meet zero entries of x_R)
      buf[u] = load_pack(base + (r < rows ? r : 0) + static_cast<int64_t>(c) * ld);
    }
  }
};

// Forward, rows [h kNarrowChunk, ...) of R of narrow front s (task (s, h)):
// gathers v on the front's columns and on those rows, y = W^T v1 (the h = 0
// task stores it, one partial sum per column), u = v_R - L_21 y stored to
// the update vector. The front's last task counts it as a finished child of
// its parent. NR right-hand sides. LDS per right-hand side: v1 (64), y (64),
// v_R (kNarrowChunk).
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void fm_task(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, int h, const TX* x,
                                        TX* uvec, TX* smem, Wait wait) {
  using NT = NarrowTile<TF>;
  static_assert(NR == 1 || NT::NB == 1, "several right-hand sides need the tile in registers");
  static_assert(128 + kNarrowChunk <= kVsLen, "staging buffer");
  const FrontShape F(a, s);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = h * kNarrowChunk, rows = max(0, min(kNarrowChunk, F.nr - r0));
  const TF* P = a.panel + a.panel_off[s];
  WTile<TF> W;
  W.load(P + F.m, F.ld, F.nc);
  RowGather G1, G2;
  TX bq[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) bq[r] = TX(0);
  if (tid < F.nc) {
    G1.pre(d, d.frow_off[s] + tid);
#pragma unroll
    for (int r = 0; r < NR; ++r) bq[r] = x[r * d.sx + F.f + tid];
  }
  if (tid < rows) G2.pre(d, d.frow_off[s] + F.nc + r0 + tid);
  wait();
  const TX* up = static_cast<const TX*>(d.upart);
  TX vr[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) vr[r] = TX(0);
  if (tid < F.nc) G1.values<NR>(d, uvec, up, bq);
  if (tid < rows) G2.values<NR>(d, uvec, up, vr);
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    TX* vs = smem + r * kVsLen;
    if (tid < 64) vs[tid] = tid < F.nc ? bq[r] : TX(0);
    vs[128 + tid] = tid < rows ? vr[r] : TX(0);
  }
  // the L_21 tile is in flight while y is formed
  NT T;
  T.base = P + F.nc + r0;
  T.ld = F.ld;
  T.rows = rows;
  T.nc = F.nc;
  T.load(0);
  __syncthreads();
  // y = W^T v1: each lane's column sums over its rows, then over the 16
  // lanes that hold the rest of those columns
  using WT = WTile<TF>;
  TX* yp = static_cast<TX*>(d.ypart) + d.ypart_off[s];
  each_rhs<NR>([&](int r) {
    TX* vs = smem + r * kVsLen;
    TX acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[k] = TX(0);
#pragma unroll
      for (int p = 0; p < WT::P; ++p)
#pragma unroll
        for (int i = 0; i < WT::N; ++i) {
          const int q = WT::row(p) + i;  // past 64 only for 16-bit factors
          acc[k] += static_cast<TX>(W.e[k][p].e[i]) * (q < 64 ? vs[q] : TX(0));
        }
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1)
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] += __shfl_xor(acc[k], off);
    if ((lane & 15) == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = WT::col(k);
        vs[64 + c] = c < F.nc ? acc[k] : TX(0);
        if (h == 0 && c < F.nc) st_dev(yp + r * d.syp + c, acc[k]);
      }
  });
  __syncthreads();
  // u = v_R - L_21 y: lanes with the same rows (l % LPC) are summed
  const int rr = T.row0();
  TX acc[NR][NT::N];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int e = 0; e < NT::N; ++e) acc[r][e] = TX(0);
#pragma unroll
  for (int b = 0; b < NT::NB; ++b) {
    if (b > 0) T.load(b);
#pragma unroll
    for (int u = 0; u < NT::B; ++u) {
      const int c = T.col(b * NT::B + u);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const TX y = smem[r * kVsLen + 64 + c];
#pragma unroll
        for (int e = 0; e < NT::N; ++e) acc[r][e] += static_cast<TX>(T.buf[u].e[e]) * y;
      }
    }
  }
#pragma unroll
  for (int off = NT::LPC; off < 64; off <<= 1)
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int e = 0; e < NT::N; ++e) acc[r][e] += __shfl_xor(acc[r][e], off);
  if (lane < NT::LPC)
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int e = 0; e < NT::N; ++e) {
        const int q = rr + e;
        if (q < rows) st_dev(uvec + r * d.su + a.rows_off[s] + r0 + q, smem[r * kVsLen + 128 + q] - acc[r][e]);
      }
  const int prev = block_publish(d.counters + kUDone * d.ns + s);
  const int* ms = d.meta + kMetaStride * s;
  if (prev == ms[kMetaNFu] - 1 && ms[kMetaParent] >= 0) add_dev(d.counters + kCDone * d.ns + ms[kMetaParent], 1);
}

// Backward, narrow front s: t = L_21^T x_R over every chunk of R, z =
// D^{-1} y - t, x1 = W z stored to x. NR right-hand sides. LDS per
// right-hand side: x_R of a chunk, later z (from 0); the waves' partial sums
// of t, later of x1 (from kNarrowChunk, 4 x 64).
template <int NR, class TF, class TD, class TX, class Wait>
__device__ __forceinline__ void bm_task(const SolveArgs<TF, TD>& a, const SolveDev& d, int s, TX* x, TX* smem,
                                        Wait wait) {
  // batches of 8 loads: the tile and one right-hand side's column sums of a
  // batch stay within the register budget
  using NT = NarrowTile<TF, 8>;
  static_assert(kNarrowChunk + 256 <= kVsLen, "staging buffer");
  const FrontShape F(a, s);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nch = max(1, (F.nr + kNarrowChunk - 1) / kNarrowChunk);
  const TF* P = a.panel + a.panel_off[s];
  NT T;
  T.base = P + F.nc;
  T.ld = F.ld;
  T.rows = min(kNarrowChunk, F.nr);
  T.nc = F.nc;
  T.load(0);
  wait();
  const int32_t* R = a.rows + a.rows_off[s];
  const int rr = T.row0();
#pragma unroll
  for (int b = 0; b < NT::NB; ++b) {
    if (b > 0 && NT::N * NT::B * b >= F.nc) break;  // no column of the front in this batch
    for (int hh = 0; hh < nch; ++hh) {
      const int rows = max(0, min(kNarrowChunk, F.nr - hh * kNarrowChunk));
      if (b > 0 || hh > 0) {
        T.base = P + F.nc + hh * kNarrowChunk;
        T.rows = rows;
        T.load(b);
      }
#pragma unroll
      for (int r = 0; r < NR; ++r)
        smem[r * kVsLen + tid] = tid < rows ? ld_dev(x + r * d.sx + R[hh * kNarrowChunk + tid]) : TX(0);
      __syncthreads();
      // the chunk's column sums of one right-hand side at a time: the lanes
      // with the same columns (l / LPC) are summed, and lane l % LPC == 0 of
      // wave w adds them to the wave's running sums (chunk order)
      each_rhs<NR>([&](int r) {
        TX acc[NT::B];
#pragma unroll
        for (int u = 0; u < NT::B; ++u) {
          acc[u] = TX(0);
#pragma unroll
          for (int e = 0; e < NT::N; ++e) acc[u] += static_cast<TX>(T.buf[u].e[e]) * smem[r * kVsLen + rr + e];
        }
#pragma unroll
        for (int off = 1; off < NT::LPC; off <<= 1)
#pragma unroll
          for (int u = 0; u < NT::B; ++u) acc[u] += __shfl_xor(acc[u], off);
        if (lane % NT::LPC == 0)
#pragma unroll
          for (int u = 0; u < NT::B; ++u) {
            TX* t = smem + r * kVsLen + kNarrowChunk + 64 * w + T.col(b * NT::B + u);
            *t = hh == 0 ? acc[u] : *t + acc[u];
          }
      });
      __syncthreads();  // the chunk's x_R is read by every wave
    }
  }
  const TX* yp = static_cast<const TX*>(d.ypart) + d.ypart_off[s];
  if (tid < 64)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      TX* zs = smem + r * kVsLen;
      const TX* red = zs + kNarrowChunk;
      const TX t = (red[tid] + red[64 + tid]) + (red[128 + tid] + red[192 + tid]);
      zs[tid] = tid < F.nc ? static_cast<TX>(a.Dinv[F.f + tid]) * ld_dev(yp + r * d.syp + tid) - t : TX(0);
    }
  // W after the tile is used up: it does not hold registers during the chunks
  using WT = WTile<TF>;
  WT W;
  W.load(P + F.m, F.ld, F.nc);
  __syncthreads();
  // x1 = W z: each lane's row sums over its 4 columns, then over the 4 lanes
  // of the wave with the same rows, then the waves in order
  each_rhs<NR>([&](int r) {
    TX* zs = smem + r * kVsLen;
    TX part[WT::P][WT::N];
#pragma unroll
    for (int p = 0; p < WT::P; ++p)
#pragma unroll
      for (int i = 0; i < WT::N; ++i) {
        part[p][i] = TX(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) part[p][i] += static_cast<TX>(W.e[k][p].e[i]) * zs[WT::col(k)];
        part[p][i] += __shfl_xor(part[p][i], 16);
        part[p][i] += __shfl_xor(part[p][i], 32);
      }
    if (lane < 16)
#pragma unroll
      for (int p = 0; p < WT::P; ++p)
#pragma unroll
        for (int i = 0; i < WT::N; ++i)
          if (WT::row(p) + i < 64) zs[kNarrowChunk + 64 * w + WT::row(p) + i] = part[p][i];
  });
  __syncthreads();
  if (tid < F.nc)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const TX* red = smem + r * kVsLen + kNarrowChunk;
      st_dev(x + r * d.sx + F.f + tid, (red[tid] + red[64 + tid]) + (red[128 + tid] + red[192 + tid]));
    }
  block_publish(d.counters + kBDone * d.ns + s);
}

// the application belongs to an iteration that has stopped (SolveDev::gate)
__device__ __forceinline__ bool gated_off(const SolveDev& d) { return d.gate != nullptr && *d.gate == 0; }

// ---------------------------------------------------------------- level schedule
template <int NR, class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_fwd_small(SolveArgs<TF, TD> a, SolveDev d, const int32_t* __restrict__ list,
                                                  TX* x, TX* uvec) {
  if (gated_off(d)) return;
  small_fwd<NR>(a, d, list[blockIdx.x], x, uvec, NoWait{});
}

template <int NR, class TF, class TD, class TX>
__global__ void __launch_bounds__(64) k_bwd_small(SolveArgs<TF, TD> a, SolveDev d, const int32_t* __restrict__ list,
                                                  TX* x) {
  __shared__ TX w[NR * kSmallFront];
  if (gated_off(d)) return;
  small_bwd<NR>(a, d, list[blockIdx.x], x, w, NoWait{});
}

// kind: kDagFx, kDagFu, kDagBr or kDagBx; tasks: the level's triples
template <int NR, class TF, class TD, class TX>
__global__ void __launch_bounds__(256) k_level_tiles(SolveArgs<TF, TD> a, SolveDev d, int kind,
                                                     const int32_t* __restrict__ tasks, TX* x, TX* uvec) {
  __shared__ TX smem[NR * (kVsLen + 4 * kRB)];
  __shared__ int flag;
  if (gated_off(d)) return;
  const int* q = tasks + 3 * blockIdx.x;
  if (kind == kDagFx) fx_task<NR>(a, d, q[0], q[1], q[2], static_cast<const TX*>(x), uvec, smem, NoWait{});
  else if (kind == kDagFu) fu_task<NR>(a, d, q[0], q[1], q[2], uvec, smem, NoWait{});
  else if (kind == kDagBr) br_task<NR>(a, d, q[0], q[1], q[2], static_cast<const TX*>(x), smem, NoWait{});
  else if (kind == kDagBx) bx_task<NR>(a, d, q[0], q[1], q[2], x, smem, &flag, NoWait{});
  else if (kind == kDagFm) fm_task<NR>(a, d, q[0], q[1], static_cast<const TX*>(x), uvec, smem, NoWait{});
  else bm_task<NR>(a, d, q[0], x, smem, NoWait{});
}

// ---------------------------------------------------------------- dag schedule
// A wait that also records, in *slot when that is set, the device clock at
// which it ended (thread 0).
template <class W>
struct Traced {
  W w;
  unsigned long long* slot;
  __device__ void operator()() const {
    w();
    if (slot != nullptr && threadIdx.x == 0) *slot = wall_clock64();
  }
};
template <class W>
__device__ __forceinline__ Traced<W> traced(W w, unsigned long long* slot) {
  return Traced<W>{w, slot};
}
// the trace record of task t (nullptr: no trace): kind and supernode, the
// clock when drawn
__device__ __forceinline__ unsigned long long* trace_begin(const SolveDev& d, int64_t t, const Task4& tk, int s) {
  if (d.trace == nullptr) return nullptr;
  unsigned long long* r = d.trace + 4 * t;
  if (threadIdx.x == 0) {
    r[0] = static_cast<unsigned long long>(tk.a) | static_cast<unsigned long long>(s) << 4;
    r[1] = wall_clock64();
  }
  return r;
}
__device__ __forceinline__ void trace_end(unsigned long long* r) {
  if (r != nullptr && threadIdx.x == 0) r[3] = wall_clock64() << 16 | blockIdx.x;
}

// One right-hand side: four waves per SIMD (at most 128 registers), 1024
// resident workgroups; several: two waves per SIMD (256 registers), so the
// factor tile and the partial sums of one right-hand side stay in registers.
template <int NR, class TF, class TD, class TX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NR == 1 ? 4 : 2))) k_fwd_dag(SolveArgs<TF, TD> a, SolveDev d, TX* x, TX* uvec) {
  __shared__ TX smem[NR * (kVsLen + 4 * kRB)];
  __shared__ int tsh, flag;
  const int w = threadIdx.x >> 6;
  int* cnt = d.counters;
  const int ns = d.ns;
  if (gated_off(d)) return;
  int* lines = cnt + d.cnt_disp;
  Dispenser ds = arrive(lines, &tsh);
  for (;;) {
    const int t = next_task(lines, ds, d.nfwd, &tsh);
    if (t >= d.nfwd) break;
    const Task4 tk = d.fwd[t];
    unsigned long long* tr = trace_begin(d, t, tk, tk.a == kDagSmall ? d.small_list[tk.b] : tk.b);
    unsigned long long* tw = tr != nullptr ? tr + 2 : nullptr;
    if (tk.a == kDagSmall) {
      if (w < tk.c) {
        const int s = d.small_list[tk.b + w];
        const int* ms = d.meta + kMetaStride * s;
        small_fwd<NR>(a, d, s, x, uvec, traced(WaveWait{cnt + kCDone * ns + s, ms[kMetaNChild], d.err}, tw));
        if (ms[kMetaParent] >= 0) wave_publish(cnt + kCDone * ns + ms[kMetaParent]);
      }
    } else if (tk.a == kDagFx) {
      const int s = tk.b;
      fx_task<NR>(a, d, s, tk.c, tk.d, static_cast<const TX*>(x), uvec, smem,
                  traced(BlockWait{cnt + kCDone * ns + s, d.meta[kMetaStride * s + kMetaNChild], d.err}, tw));
    } else if (tk.a == kDagFu) {
      const int s = tk.b;
      fu_task<NR>(a, d, s, tk.c, tk.d, uvec, smem,
                  traced(BlockWait{cnt + kXDone * ns + s, d.meta[kMetaStride * s + kMetaNFx], d.err}, tw));
    } else {  // kDagFm
      const int s = tk.b;
      fm_task<NR>(a, d, s, tk.c, static_cast<const TX*>(x), uvec, smem,
                  traced(BlockWait{cnt + kCDone * ns + s, d.meta[kMetaStride * s + kMetaNChild], d.err}, tw));
    }
    trace_end(tr);
  }
}

template <int NR, class TF, class TD, class TX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NR == 1 ? 4 : 2))) k_bwd_dag(SolveArgs<TF, TD> a, SolveDev d, TX* x) {
  __shared__ TX smem[NR * (kVsLen + 4 * kRB)];
  __shared__ int tsh, flag;
  const int w = threadIdx.x >> 6;
  int* cnt = d.counters;
  const int ns = d.ns;
  if (gated_off(d)) return;
  int* lines = cnt + d.cnt_disp + kDispenserLines * kDispenserStride;
  Dispenser ds = arrive(lines, &tsh);
  for (;;) {
    const int t = next_task(lines, ds, d.nbwd, &tsh);
    if (t >= d.nbwd) break;
    const Task4 tk = d.bwd[t];
    unsigned long long* tr =
        trace_begin(d, static_cast<int64_t>(d.nfwd) + t, tk, tk.a == kDagSmall ? d.small_list[tk.b] : tk.b);
    unsigned long long* tw = tr != nullptr ? tr + 2 : nullptr;
    if (tk.a == kDagSmall) {
      if (w < tk.c) {
        const int s = d.small_list[tk.b + w];
        const int p = d.meta[kMetaStride * s + kMetaParent];
        // a root waits for nothing: its target is 0
        const int* pc = cnt + kBDone * ns + (p >= 0 ? p : s);
        const int target = p >= 0 ? d.meta[kMetaStride * p + kMetaNBFinal] : 0;
        small_bwd<NR>(a, d, s, x, smem + kRB * NR * w, traced(WaveWait{pc, target, d.err}, tw));
        wave_publish(cnt + kBDone * ns + s);
      }
    } else if (tk.a == kDagBr) {
      const int s = tk.b;
      const int p = d.meta[kMetaStride * s + kMetaParent];
      const int* pc = cnt + kBDone * ns + (p >= 0 ? p : s);
      const int target = p >= 0 ? d.meta[kMetaStride * p + kMetaNBFinal] : 0;
      br_task<NR>(a, d, s, tk.c, tk.d, static_cast<const TX*>(x), smem, traced(BlockWait{pc, target, d.err}, tw));
    } else if (tk.a == kDagBx) {
      const int s = tk.b;
      bx_task<NR>(a, d, s, tk.c, tk.d, x, smem, &flag,
                  traced(BlockWait{cnt + kRDone * ns + s, d.meta[kMetaStride * s + kMetaNBr], d.err}, tw));
    } else {  // kDagBm
      const int s = tk.b;
      const int p = d.meta[kMetaStride * s + kMetaParent];
      const int* pc = cnt + kBDone * ns + (p >= 0 ? p : s);
      const int target = p >= 0 ? d.meta[kMetaStride * p + kMetaNBFinal] : 0;
      bm_task<NR>(a, d, s, x, smem, traced(BlockWait{pc, target, d.err}, tw));
    }
    trace_end(tr);
  }
}

__global__ void k_zero(int n, int* __restrict__ p, const int* __restrict__ gate) {
  if (gate != nullptr && *gate == 0) return;
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

// NR right-hand sides x + r sx (r < NR) in one pair of launches
template <int NR, class TF, class TD, class TX>
void ldl_solve_group(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, const SolveDev& d, TX* x, TX* uvec,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3((d.ncounters + 255) / 256), dim3(256), 0, st, d.ncounters, d.counters, d.gate);
  if (dp.schedule == kScheduleDag) {
    // at least kDispensers workgroups, so that every class can be covered
    // (next_task states why the persistent launch always progresses)
    static const int grid_f = resident_blocks(k_fwd_dag<NR, TF, TD, TX>, 256);
    static const int grid_b = resident_blocks(k_bwd_dag<NR, TF, TD, TX>, 256);
    if (d.nfwd > 0)
      hipLaunchKernelGGL((k_fwd_dag<NR, TF, TD, TX>), dim3(std::max(kDispensers, std::min(d.nfwd, grid_f))), dim3(256),
                         0, st, a, d, x, uvec);
    if (d.nbwd > 0)
      hipLaunchKernelGGL((k_bwd_dag<NR, TF, TD, TX>), dim3(std::max(kDispensers, std::min(d.nbwd, grid_b))), dim3(256),
                         0, st, a, d, x);
    MPIR_HIP(hipGetLastError());
    return;
  }
  auto tiles = [&](int kind, const std::vector<int32_t>& ptr, const int32_t* list, int l) {
    const int t0 = ptr[l], t1 = ptr[l + 1];
    if (t1 > t0)
      hipLaunchKernelGGL((k_level_tiles<NR, TF, TD, TX>), dim3(t1 - t0), dim3(256), 0, st, a, d, kind, list + 3 * t0,
                         x, uvec);
  };
  // forward, leaves to root; D^{-1} fused
  for (int l = 0; l < dp.nlevels; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_fwd_small<NR, TF, TD, TX>), dim3(s1 - s0), dim3(64), 0, st, a, d, d.small_list + s0, x,
                         uvec);
    tiles(kDagFx, dp.h_fx_ptr, d.fx_task, l);
    tiles(kDagFu, dp.h_fu_ptr, d.fu_task, l);
    tiles(kDagFm, dp.h_fm_ptr, d.fm_task, l);
  }
  // backward, root to leaves
  for (int l = dp.nlevels - 1; l >= 0; --l) {
    tiles(kDagBr, dp.h_br_ptr, d.br_task, l);
    tiles(kDagBx, dp.h_bx_ptr, d.bx_task, l);
    tiles(kDagBm, dp.h_bm_ptr, d.bm_task, l);
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_bwd_small<NR, TF, TD, TX>), dim3(s1 - s0), dim3(64), 0, st, a, d, d.small_list + s0, x);
  }
  MPIR_HIP(hipGetLastError());
}

}  // namespace

template <class TF, class TD, class TX>
void ldl_solve(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, TX* uvec, hipStream_t st) {
  ldl_solve_group<1>(a, dp, dp.d, x, uvec, st);
}

template <class TF, class TD, class TX>
void ldl_solve_on(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, const SolveDev& d, TX* x, TX* uvec,
                  hipStream_t st) {
  ldl_solve_group<1>(a, dp, d, x, uvec, st);
}

template <class TF, class TD, class TX>
void ldl_solve_multi(const SolveArgs<TF, TD>& a, const DeviceSolvePlan& dp, TX* x, int nrhs, TX* uvec,
                     hipStream_t st, const int* gate) {
  SolveDev d = dp.d;
  d.gate = gate;
  // groups of 8, 4, 2, 1 right-hand sides for FP32 factors; FP64 tiles do
  // not fit in registers and 16-bit tiles with their conversions spill, so
  // those factors take the right-hand sides one at a time
  constexpr bool multi = std::is_same<TF, float>::value;
  static_assert(!multi || (ColTile<TF>::NB == 1 && RowTile<TF>::NB == 1), "tile in registers");
  for (int r0 = 0; r0 < nrhs;) {
    const int left = nrhs - r0;
    TX* xr = x + r0 * d.sx;
    if constexpr (multi) {
      if (left >= 8) {
        ldl_solve_group<8>(a, dp, d, xr, uvec, st);
        r0 += 8;
        continue;
      }
      if (left >= 4) {
        ldl_solve_group<4>(a, dp, d, xr, uvec, st);
        r0 += 4;
        continue;
      }
      if (left >= 2) {
        ldl_solve_group<2>(a, dp, d, xr, uvec, st);
        r0 += 2;
        continue;
      }
    }
    ldl_solve_group<1>(a, dp, d, xr, uvec, st);
    r0 += 1;
  }
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
template void ldl_solve_on<double, double, double>(const SolveArgs<double, double>&, const DeviceSolvePlan&,
                                                    const SolveDev&, double*, double*, hipStream_t);
template void ldl_solve_on<float, float, double>(const SolveArgs<float, float>&, const DeviceSolvePlan&,
                                                  const SolveDev&, double*, double*, hipStream_t);
template void ldl_solve_on<__half, float, double>(const SolveArgs<__half, float>&, const DeviceSolvePlan&,
                                                   const SolveDev&, double*, double*, hipStream_t);
template void ldl_solve_on<__hip_bfloat16, float, double>(const SolveArgs<__hip_bfloat16, float>&,
                                                           const DeviceSolvePlan&, const SolveDev&, double*, double*,
                                                           hipStream_t);
template void ldl_solve_multi<double, double, double>(const SolveArgs<double, double>&, const DeviceSolvePlan&,
                                                       double*, int, double*, hipStream_t, const int*);
template void ldl_solve_multi<float, float, double>(const SolveArgs<float, float>&, const DeviceSolvePlan&, double*,
                                                     int, double*, hipStream_t, const int*);
template void ldl_solve_multi<__half, float, double>(const SolveArgs<__half, float>&, const DeviceSolvePlan&, double*,
                                                      int, double*, hipStream_t, const int*);
template void ldl_solve_multi<__hip_bfloat16, float, double>(const SolveArgs<__hip_bfloat16, float>&,
                                                              const DeviceSolvePlan&, double*, int, double*,
                                                              hipStream_t, const int*);

}  // namespace mpir
