// Multifrontal LDL^T for quasi-definite matrices on gfx950.
//
// Numerically this is QDLDL_factor (include/MPIR/qdldl.hpp:172-345 of
// yex33/MPIR): no pivoting, L unit lower triangular, D diagonal with
// Dinv = 1 / D (qdldl.hpp:315), each L entry a scaled Schur complement entry
// (qdldl.hpp:339: y * Dinv; small fronts do exactly that, big fronts form the
// rows below a diagonal block as products with its explicit inverse and the
// block's own multipliers with a Newton-refined reciprocal, recip_nr), and a
// zero pivot reported as failure (qdldl.hpp:331-333). The loop order is
// different: columns are grouped into supernodes and eliminated front by
// front, level by level of the supernodal tree, which is what exposes
// parallelism.
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
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "../common.hpp"
#include "../plan.hpp"
#include "kernels.hpp"
#include "sync.hpp"

namespace mpir {

namespace {

__device__ __forceinline__ void flag_zero_pivot(int* err) { atomicOr(err, 1); }

// Memory accesses of the task bodies: plain for the per-level launches
// (kernel boundaries order them), device-coherent inside the
// dependency-driven launch, where another workgroup of the same launch
// wrote the values.
struct PlainMem {
  template <class T>
  static __device__ __forceinline__ T ld(const T* p) {
    return *p;
  }
  template <class T>
  static __device__ __forceinline__ void st(T* p, T v) {
    *p = v;
  }
};
// 16 bytes of consecutive entries: loaded from any address with the entry's
// own alignment, stored to LDS at 16-byte aligned places
template <class T>
struct alignas(16) Pack16 {
  static constexpr int N = 16 / sizeof(T);
  T e[N];
};
template <class T>
__device__ __forceinline__ Pack16<T> load_pack16(const T* p) {
  Pack16<T> r;
  __builtin_memcpy(&r, p, sizeof r);
  return r;
}

struct CoherentMem {
  template <class T>
  static __device__ __forceinline__ T ld(const T* p) {
    return ld_dev(p);
  }
  template <class T>
  static __device__ __forceinline__ void st(T* p, T v) {
    st_dev(p, v);
  }
};

// Rounding of the factor's entries when the factor is held in a 16-bit
// format: R = 0 none, 1 to FP16, 2 to BF16 (round to nearest even). The
// matrix values entering the factorization and every entry of L are rounded
// as they are stored, so the trailing updates multiply 16-bit operands;
// products, sums and the update matrices (Schur complements) stay in T
// (FP32 accumulation). D, D^{-1} and the inverse rows are not rounded.
template <int R, class T>
__device__ __forceinline__ T rnd(T v) {
  if constexpr (R == 1) return static_cast<T>(__half2float(__float2half_rn(static_cast<float>(v))));
  else if constexpr (R == 2) return static_cast<T>(__bfloat162float(__float2bfloat16(static_cast<float>(v))));
  else return v;
}

// 1 / d from the hardware reciprocal estimate refined by Newton-Raphson
// steps x <- x + x (1 - d x) (one in FP32, two in FP64): within an ulp or so
// of the correctly rounded quotient, at a fraction of a division's
// dependent latency. The diagonal-block eliminations of big fronts form
// their multipliers with it (the pivots of a block are a serial chain);
// D^{-1} itself is stored as the quotient 1 / d.
template <class T>
__device__ __forceinline__ T recip_nr(T d) {
  if constexpr (sizeof(T) == 4) {
    const float x = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, x, 1.f), x, x);
  } else {
    double x = __builtin_amdgcn_rcp(d);
    x = __builtin_fma(__builtin_fma(-d, x, 1.0), x, x);
    return __builtin_fma(__builtin_fma(-d, x, 1.0), x, x);
  }
}

// ---------------------------------------------------------------- scatter
template <class T, int R>
__global__ void k_scatter_values(int64_t nnz, const double* __restrict__ ax,
                                 const int64_t* __restrict__ dst, T* __restrict__ panel) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < nnz) panel[dst[q]] = rnd<R>(static_cast<T>(ax[q]));
}

// ---------------------------------------------------------------- small fronts
// One wavefront per front (m <= 64), the whole front in LDS.
template <class T, int R>
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

  // load the panel columns, zero the update block (loads unconditional from
  // clamped addresses so that they are all in flight together)
  {
    const int ls = min(lane, m - 1);
#pragma unroll 16
    for (int c = 0; c < kSmallFront; ++c) {
      const T v = P[ls + static_cast<int64_t>(min(c, nc - 1)) * m];
      F[lane][c] = (lane < m && c < nc) ? v : T(0);
    }
  }
  __syncthreads();
  // extend-add the children's update matrices, one child at a time
  for (int t = a.child_ptr[s]; t < a.child_ptr[s + 1]; ++t) {
    const int c = a.child_list[t];
    const int nrc = a.nrow[c];
    if (nrc == 0) continue;
    const T* Uc = a.upool + a.upd_off[c];
    const int32_t* rel = a.relidx + a.rows_off[c];
    const int ls = min(lane, nrc - 1);
    const int ra = rel[ls];
    for (int b0 = 0; b0 < nrc; b0 += 8) {
      T uv[8];
      int rb[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int b = min(b0 + q, nrc - 1);
        rb[q] = rel[b];
        uv[q] = Uc[ls + static_cast<int64_t>(b) * nrc];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int b = b0 + q;
        if (b < nrc && lane >= b && lane < nrc) F[ra][rb[q]] += uv[q];
      }
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
    __syncthreads();
  }
  if (lane < nc) {
    const T d = F[lane][lane];
    a.D[f + lane] = d;
    a.Dinv[f + lane] = T(1) / d;
  }
  for (int c = 0; c < nc; ++c)
    if (lane < m && lane >= c) P[lane + static_cast<int64_t>(c) * m] = lane == c ? F[lane][c] : rnd<R>(F[lane][c]);
  for (int b = 0; b < nr; ++b)
    if (lane >= b && lane < nr) U[lane + static_cast<int64_t>(b) * nr] = F[nc + lane][nc + b];
}

// ---------------------------------------------------------------- assembly
// Tile (s, c0, c1, t0): columns [c0, c1) and rows [t0, t0 + 64) of front s
// (lower part). The tile is accumulated in LDS: panel columns start from the
// scattered matrix entries, update-matrix columns from zero; the children's
// update matrices are added in fixed order (a barrier between children, no
// two entries of one child land on the same place), then the tile is
// written back once.
// Children entries [e0, e1) of cl; tile: E x (E + 1) entries of LDS. A
// child's entries are loaded KB per thread at a time (no two land on the same
// place, so the rounds do not change any sum).
template <class T, int R, class M, int KB = kAsmTile * kAsmTile / 256>
__device__ __forceinline__ void asm_body(const FactorArgs<T>& a, const Task4 tk, int e0, int e1,
                                         const int32_t* __restrict__ cl, T (*tile)[kAsmTile + 1]) {
  constexpr int E = kAsmTile;
  const int s = tk.a, c0 = tk.b, c1 = tk.c, t0 = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int t1 = min(m, t0 + E);
  T* P = a.panel + a.panel_off[s];
  T* U = nr > 0 ? a.upool + a.upd_off[s] : P;
  const int tid = threadIdx.x, i = tid & 63, w = tid >> 6;
  const int r = t0 + i;
  if constexpr (std::is_same<M, PlainMem>::value) {
    // 16-byte packs of PN rows of one panel column: thread (rows pi ..,
    // columns pt + PC j); a pack may run up to PN - 1 rows past the tile,
    // inside the column or the pool's zeroed slack (those rows are not kept)
    using Pk = Pack16<T>;
    constexpr int PN = Pk::N, PL = E / PN, PC = 256 / PL, NP = E / PC;
    const int pi = PN * (tid % PL), pt = tid / PL;
    Pk v[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int c = c0 + pt + PC * j;
      v[j] = load_pack16(c < c1 && c < nc ? P + t0 + pi + c * ld : P);
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int c = c0 + pt + PC * j;
#pragma unroll
      for (int e = 0; e < PN; ++e) {
        const int re = t0 + pi + e;
        const bool in = c < c1 && re < t1 && re >= c && c < nc;
        tile[pi + e][pt + PC * j] = in ? v[j].e[e] : T(0);
      }
    }
  } else {
    T v[E / 4];
#pragma unroll
    for (int q = 0; q < E / 4; ++q) {
      const int c = c0 + w + 4 * q;
      const bool in = c < c1 && r < t1 && r >= c && c < nc;
      v[q] = M::ld(in ? P + r + c * ld : P);
      tile[i][w + 4 * q] = in ? v[q] : T(0);
    }
  }
  __syncthreads();
  if constexpr (std::is_same<M, PlainMem>::value) {
    // A child's block is at most E rows x E columns (its rows and columns
    // map into the tile's, strictly increasing): thread (rows a0 + pi ..,
    // columns b0 + pt + PC j) reads each column's PN rows as one 16-byte
    // pack (up to PN - 1 rows past the block: inside the pool, which ends in
    // slack; not used).
    using Pk = Pack16<T>;
    constexpr int PN = Pk::N, PL = E / PN, PC = 256 / PL, NP = E / PC;
    const int pi = PN * (tid % PL), pt = tid / PL;
    for (int e = e0; e < e1; ++e) {
      const int ch = cl[5 * e], b0 = cl[5 * e + 1], b1 = cl[5 * e + 2], a0 = cl[5 * e + 3], a1 = cl[5 * e + 4];
      const int nrc = a.nrow[ch];
      const T* Uc = a.upool + a.upd_off[ch];
      const int32_t* rel = a.relidx + a.rows_off[ch];
      const int q0 = a0 + pi;
      Pk v[NP];
      int dc[NP], bq[NP];
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int b = b0 + pt + PC * j;
        const int bs = b < b1 ? b : b0;
        v[j] = load_pack16(Uc + (q0 < a1 ? q0 : a0) + static_cast<int64_t>(bs) * nrc);
        dc[j] = rel[bs] - c0;
        bq[j] = b < b1 ? b : nrc;  // a column past the block takes no row
      }
      int dr[PN];
#pragma unroll
      for (int u = 0; u < PN; ++u) dr[u] = q0 + u < a1 ? rel[min(q0 + u, a1 - 1)] - t0 : -1;
#pragma unroll
      for (int j = 0; j < NP; ++j)
#pragma unroll
        for (int u = 0; u < PN; ++u)
          if (dr[u] >= 0 && q0 + u >= bq[j]) tile[dr[u]][dc[j]] += v[j].e[u];
      __syncthreads();
    }
  } else {
    for (int e = e0; e < e1; ++e) {
      const int ch = cl[5 * e], b0 = cl[5 * e + 1], b1 = cl[5 * e + 2], a0 = cl[5 * e + 3], a1 = cl[5 * e + 4];
      const int nrc = a.nrow[ch];
      const T* Uc = a.upool + a.upd_off[ch];
      const int32_t* rel = a.relidx + a.rows_off[ch];
      const int na = a1 - a0, npair = (b1 - b0) * na;
      static_assert((E * E / 256) % KB == 0, "assembly rounds");
#pragma unroll
      for (int k0 = 0; k0 < E * E / 256; k0 += KB) {
        T v[KB];
        int dr[KB], dc[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          const int p = tid + 256 * (k0 + k);
          const int b = b0 + (p < npair ? p / na : 0);
          const int q = a0 + (p < npair ? p % na : 0);
          v[k] = M::ld(Uc + q + static_cast<int64_t>(b) * nrc);
          dr[k] = (p < npair && q >= b) ? rel[q] - t0 : -1;
          dc[k] = rel[b] - c0;
        }
#pragma unroll
        for (int k = 0; k < KB; ++k)
          if (dr[k] >= 0) tile[dr[k]][dc[k]] += v[k];
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < E / 4; ++q) {
    const int c = c0 + w + 4 * q;
    if (c < c1 && r < t1 && r >= c) {
      if (c < nc) M::st(P + r + c * ld, tile[i][w + 4 * q]);
      else M::st(U + (r - nc) + static_cast<int64_t>(c - nc) * nr, tile[i][w + 4 * q]);
    }
  }
}

template <class T, int R>
__global__ void __launch_bounds__(256) k_assemble(FactorArgs<T> a, const Task4* __restrict__ tasks,
                                                  const int32_t* __restrict__ cptr, const int32_t* __restrict__ cl) {
  __shared__ T tile[kAsmTile][kAsmTile + 1];  // tile[row - t0][col - c0]
  asm_body<T, R, PlainMem>(a, tasks[blockIdx.x], cptr[blockIdx.x], cptr[blockIdx.x + 1], cl, tile);
}

// Matrix-core product for the trailing updates: v_mfma 16x16x4 in the
// precision of the factor. Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; output v (0..3) of lane l is D[row(l / 16, v)][l % 16]
// with row = 4 (l / 16) + v for fp32 and (l / 16) + 4 v for fp64 (measured
// on gfx950).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <class T>
struct Mfma16;
template <>
struct Mfma16<float> {
  using V = f32x4;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};
template <>
struct Mfma16<double> {
  using V = f64x4;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};

constexpr int kXcds = 8;  // accelerator complex dies of an MI355X, each with its own L2
// task of workgroup b when workgroup b runs on XCD b % kXcds: XCD x takes the
// x-th contiguous run of the list (runs differ in length by at most one)
__device__ __forceinline__ int xcd_task(int b, int n) {
  const int x = b % kXcds, i = b / kXcds;
  const int q = n / kXcds, rem = n % kXcds;
  return x * q + min(x, rem) + i;
}

// ---------------------------------------------------------------- big fronts
// Panel step k of big front s, task (s, r0, r1, k). Every workgroup factors
// the nbk x nbk diagonal block in LDS (right-looking LDL^T) and carries 64
// identity rows through the same elimination, which leaves
// Z = L_kk^{-T} D_k^{-1} (upper triangular) in them. The task with
// r0 = k0 owns the diagonal block and writes L_kk, D and D^{-1}; every other
// task forms its rows of L as X Z, X = its rows of the current panel
// (identity rows of the front, r >= m, enter at their own column as e_i).
// refactor (template) = false: L_kk and D are already final (a factor loaded from
// outside); the block is rebuilt from them and only Z is formed.
// PM (launch split, refactor only): 0 every task as above; 1 only the tasks
// that own a diagonal block, which also store Z_kk to the front's slot of
// a.zbuf; 2 only the other tasks, with Z_kk read from that slot instead of
// eliminated again (the same values: every task of a front eliminates the
// same block with the same operations).
constexpr int kPanelThreads = 1024;  // 16 waves, each owning 4 block columns

template <class T, int R, bool refactor, int PM = 0>
__global__ void __launch_bounds__(kPanelThreads) k_panel_step(FactorArgs<T> a, const Task4* __restrict__ tasks) {
  static_assert(PM == 0 || refactor, "split launches refactor");
  constexpr int NB = kPanelNB;
  constexpr int CW = NB / (kPanelThreads / 64);  // columns per wave (4)
  // per block step J: the multipliers of every row of the block (mrs), of
  // every identity row (gms) and d_t times the multipliers of every column
  // (dms), double-buffered
  __shared__ __align__(16) T mrs[2][NB][CW];
  __shared__ __align__(16) T gms[2][NB][CW];
  __shared__ __align__(16) T dms[2][NB][CW];
  // row stride NB + 1: the column-wise accesses (Zs[lane][c], Xs[lane][rr])
  // are conflict-free
  __shared__ __align__(16) T Zs[NB][NB + 1];  // Z, row-major
  __shared__ __align__(16) T Xs[NB][NB + 1];  // the task's rows, transposed: Xs[t][row]
  const Task4 tk = tasks[blockIdx.x];
  const int s = tk.a, r0 = tk.b, r1 = tk.c, k = tk.d;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int64_t ld = panel_ld(nc, a.nrow[s]);
  const int k0 = k * NB;
  const int nbk = min(NB, nc - k0);
  T* P = a.panel + a.panel_off[s];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c0w = CW * w;  // this thread's columns: c0w .. c0w + CW - 1
  const bool owner = r0 == k0;
  if (PM == 1 && !owner) return;
  if (PM == 2 && owner) return;
  const int rows = r1 - r0;
  // Loads are unconditional from clamped addresses, then selected: a load
  // under a branch is waited for before the branch joins.
  const int r = r0 + lane;
  const int unit = r >= m + k0 ? r - m - k0 : -1;
  const bool live = !owner && lane < rows;
  const int64_t rsafe = (live && unit < 0) ? r : k0;
  T xv[CW];
#pragma unroll
  for (int u = 0; u < CW; ++u) {
    const int c = c0w + u;
    const T v = P[rsafe + (k0 + min(c, nbk - 1)) * ld];
    xv[u] = (!live || c >= nbk) ? T(0) : unit >= 0 ? (c == unit ? T(1) : T(0)) : v;
  }
  // Entries (lane, c0w + u) of the block's lower triangle (fa) and of the
  // identity rows (ga), in registers through the elimination. With refactor
  // fa holds the block as eliminated so far; otherwise it holds the given
  // factor (L below the diagonal, from its transposed slot, and D on it).
  const T* Dk = a.D + f + k0;
  const int lsafe = min(lane, nbk - 1);
  T fa[CW], ga[CW];
  // the front's Z_kk slot (split launches): entry (row i, column c) at
  // i + kPanelNB c
  const int zi = PM != 0 ? a.zslot[s] : 0;
  if (PM != 0 && (zi < 0 || zi >= a.zcap)) {  // no slot: reported, nothing stored
    if (tid == 0) atomicOr(a.err, 4);
    return;
  }
  T* zs = PM != 0 ? a.zbuf + static_cast<int64_t>(zi) * (NB * NB) : nullptr;
  if constexpr (PM == 2) {
#pragma unroll
    for (int u = 0; u < CW; ++u) ga[u] = zs[lane + (c0w + u) * NB];
  } else {
#pragma unroll
    for (int u = 0; u < CW; ++u) {
      const int c = c0w + u;
      const int cs = min(c, nbk - 1);
      const T v = refactor ? P[(k0 + lsafe) + (k0 + cs) * ld] : P[(k0 + cs) + (k0 + lsafe) * ld];
      const T dc = Dk[cs];
      const bool in = lane < nbk && c < nbk && lane >= c;
      fa[u] = !in ? T(0) : (refactor || lane != c) ? v : dc;
      ga[u] = (lane == c && c < nbk) ? T(1) : T(0);
    }
  }
  // Blocked right-looking elimination, CW columns per barrier. Wave J owns
  // columns CW J .. CW J + CW - 1; when they are final it factors the CW x CW
  // pivot block (read across its lanes), forms the multipliers of every row
  // (its lanes are the rows) and of every identity row, and publishes them
  // with d_t times the multipliers of every column, which by symmetry are
  // those of the row of the same index. After the barrier the waves right of
  // it apply the rank-CW update to their columns.
  const int nstep = PM == 2 ? 0 : (nbk + CW - 1) / CW;
  for (int J = 0; J < nstep; ++J) {
    const int b = J & 1;
    const int j0 = CW * J;
    if (w == J) {
      bool valid[CW];
#pragma unroll
      for (int t = 0; t < CW; ++t) valid[t] = j0 + t < nbk;
      // pivot block L_B D_B L_B^T; dl[t][q] = d_q l_tq, dinv = 1 / d
      T d[CW], dinv[CW], l[CW][CW], dl[CW][CW];
#pragma unroll
      for (int t = 0; t < CW; ++t) {
        T dt = lane_value(fa[t], j0 + t);
        if (refactor) {
#pragma unroll
          for (int q = 0; q < t; ++q) dt -= l[t][q] * dl[t][q];
        }
        d[t] = valid[t] ? dt : T(1);
        dinv[t] = recip_nr(d[t]);
#pragma unroll
        for (int u = t + 1; u < CW; ++u) {
          T v = lane_value(fa[t], min(j0 + u, NB - 1));
          if (refactor) {
#pragma unroll
            for (int q = 0; q < t; ++q) v -= l[u][q] * dl[t][q];
            v *= dinv[t];
          }
          l[u][t] = (valid[u] && valid[t]) ? v : T(0);
          dl[u][t] = d[t] * l[u][t];
        }
      }
      // multipliers of this lane's row and identity row
      T mr[CW], gm[CW];
#pragma unroll
      for (int t = 0; t < CW; ++t) {
        T v = fa[t];
        T g = ga[t];
#pragma unroll
        for (int q = 0; q < t; ++q) {
          if (refactor) v -= mr[q] * dl[t][q];
          g -= gm[q] * dl[t][q];
        }
        mr[t] = refactor ? v * dinv[t] : v;
        gm[t] = g * dinv[t];
      }
#pragma unroll
      for (int t = 0; t < CW; ++t) {
        mrs[b][lane][t] = mr[t];
        gms[b][lane][t] = gm[t];
        dms[b][lane][t] = d[t] * mr[t];
      }
      // the owner keeps its columns of L_kk (and D on the diagonal) in LDS;
      // the identity rows' entries in the block columns are final
      if (owner) {
        const int s0 = lane - j0;
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          T lv = T(0);
#pragma unroll
          for (int u = t + 1; u < CW; ++u) lv = s0 == u ? l[u][t] : lv;
          Xs[j0 + t][lane] = s0 < t ? T(0) : s0 == t ? d[t] : s0 < CW ? lv : mr[t];
        }
      }
#pragma unroll
      for (int t = 0; t < CW; ++t)
        if (valid[t]) ga[t] = gm[t];
    }
    __syncthreads();
    // rank-CW update of this thread's entries right of the block
    if (w > J) {
      // every operand is read before any is used (no loads under a branch)
      T mr[CW], gm[CW], dm[CW][CW];
#pragma unroll
      for (int t = 0; t < CW; ++t) {
        mr[t] = mrs[b][lane][t];
        gm[t] = gms[b][lane][t];
      }
#pragma unroll
      for (int u = 0; u < CW; ++u)
#pragma unroll
        for (int t = 0; t < CW; ++t) dm[u][t] = dms[b][c0w + u][t];
#pragma unroll
      for (int u = 0; u < CW; ++u) {
        const int c = c0w + u;
        T fu = T(0), gu = T(0);
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          fu += mr[t] * dm[u][t];
          gu += gm[t] * dm[u][t];
        }
        const bool on = c < nbk;
        fa[u] = (on && refactor && lane >= c) ? fa[u] - fu : fa[u];
        ga[u] = on ? ga[u] - gu : ga[u];
      }
    }
  }
  if (owner) {
    if (!refactor) return;
    if constexpr (PM == 1) {
#pragma unroll
      for (int u = 0; u < CW; ++u) zs[lane + (c0w + u) * NB] = ga[u];
    }
    __syncthreads();
    // L_kk goes to the block's upper triangle, transposed (entry (j, r) holds
    // L[r, j]): the other workgroups of this launch read the lower triangle
    for (int rr = w; rr < nbk; rr += kPanelThreads / 64)
      if (lane < rr) P[(k0 + lane) + (k0 + rr) * ld] = rnd<R>(Xs[lane][rr]);
    if (tid < nbk) {
      const T d = Xs[tid][tid];
      if (d == T(0)) flag_zero_pivot(a.err);
      a.D[f + k0 + tid] = d;
      a.Dinv[f + k0 + tid] = T(1) / d;
    }
    return;
  }
  // rows of L: X Z on the matrix cores, wave w the 16 x 16 block (rows
  // 16 (w % 4) .., columns 16 (w / 4) ..); rows and columns t >= nbk of Xs
  // and Zs are zero. The product goes through LDS (Xs, as [column][row]) to
  // thread (row lane, columns c0w ..).
#pragma unroll
  for (int u = 0; u < CW; ++u) {
    Zs[lane][c0w + u] = ga[u];
    Xs[c0w + u][lane] = xv[u];
  }
  __syncthreads();
  using V = typename Mfma16<T>::V;
  const int li = lane & 15, lk = lane >> 4;
  const int bx = w & 3, by = w >> 2;
  V o;
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = T(0);
  for (int t0 = 0; t0 < nbk; t0 += 4)
    o = Mfma16<T>::op(Xs[t0 + lk][16 * bx + li], Zs[t0 + lk][16 * by + li], o);
  __syncthreads();
#pragma unroll
  for (int v = 0; v < 4; ++v) Xs[16 * by + li][16 * bx + Mfma16<T>::row(lk, v)] = o[v];
  __syncthreads();
  if (lane >= rows) return;
#pragma unroll
  for (int u = 0; u < CW; ++u) {
    const int c = c0w + u;
    const T acc = Xs[c][lane];
    if (c < nbk) P[r + (k0 + c) * ld] = r < m ? rnd<R>(acc) : acc;
  }
}

// Trailing update tile: C[r0:r0+64, c0:c0+64] -= sum over the task's steps
// of L_r D L_c^T (the nbk columns of each step), lower part only. A task
// applies `steps` consecutive panel steps ending at k (field d = k | steps <<
// kUpdStepsShift), accumulating them on the matrix cores before one
// read-modify-write of the target. Columns < nc live in the panel, columns
// >= nc in the update matrix. Tiles with r0 >= m cover identity rows (rows
// m .. m+nc) and only panel columns. Each wave computes a 32 x 32 quarter
// with 2 x 2 MFMA blocks; the next step's operands are loaded into registers
// while the current step's products run.
// Operand tiles Ls[0][t][i] = L[r0 + i, k0 + t] and Ls[1][t][j] =
// L[c0 + j, k0 + t] D[k0 + t], stored at column i ^ 16 (t mod 4): the four
// 16-lane groups of an MFMA operand read (rows t0 .. t0 + 3) fall in
// distinct LDS banks without padding (32 KB in FP32: five workgroups per CU).
__device__ __forceinline__ int op_at(int t, int i) { return t * kTile + (i ^ ((t & 3) << 4)); }

// Operand chunk width (k entries staged at a time) of the update tiles: 32
// for the per-level launches (16.6 KB of LDS, the staging tile included,
// so that more workgroups fit on a CU), 64 for the dependency-driven launch.
template <class M>
constexpr int upd_chunk() {
  return std::is_same<M, PlainMem>::value ? 32 : kPanelNB;
}
// LDS of the update tiles: two operand chunks, or the kTile x (kTile + 1)
// staging tile, whichever is larger
template <class M>
constexpr int upd_lds() {
  return 2 * upd_chunk<M>() * kTile > kTile * (kTile + 1) ? 2 * upd_chunk<M>() * kTile : kTile * (kTile + 1);
}

// Ls: upd_lds<M>() entries of LDS (16-byte aligned).
template <class T, int R, class M>
__device__ __forceinline__ void upd_body(const FactorArgs<T>& a, const Task4 tk, T* Ls) {
  using V = typename Mfma16<T>::V;
  auto at = [](int t, int i) { return op_at(t, i); };
  const int s = tk.a, r0 = tk.b, c0 = tk.c;
  const int klast = tk.d & ((1 << kUpdStepsShift) - 1);
  const int kfirst = klast - ((tk.d >> kUpdStepsShift) & ((1 << (kUpdPanelBit - kUpdStepsShift)) - 1)) + 1;
  const bool panel_only = (tk.d >> kUpdPanelBit) & 1;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int rlim = r0 >= m ? m + nc : m;
  const T* P = a.panel + a.panel_off[s];
  T* Pw = a.panel + a.panel_off[s];
  T* U = nr > 0 ? a.upool + a.upd_off[s] : Pw;
  // w is wave-uniform (so are the D loads of the scalar path: scalar loads)
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // The steps' k range is staged KC columns at a time (chunk q: panel
  // columns q KC ..), in order, so the products run over k in the same
  // order for either chunk width. Operand staging: all loads of a chunk
  // issued before any is used (clamped addresses, values selected
  // afterwards). Per-level launches (PlainMem) move 16-byte packs of PN
  // consecutive rows of one column: thread (pack row i = PN (tid % PL),
  // columns t = tid / PL + PC j), which fills the operand tiles with the
  // same values as the entry-wise path of the dependency-driven launch,
  // whose loads must be device-coherent. A pack starting inside the tile
  // may run up to PN - 1 rows past its last row: those rows lie inside the
  // column or in the pool's zeroed slack and are stored as zeros.
  constexpr bool packed = std::is_same<M, PlainMem>::value;
  constexpr int KC = upd_chunk<M>();
  static_assert(kPanelNB % KC == 0 && KC % 4 == 0, "chunk width");
  T* L0 = Ls;
  T* L1 = Ls + KC * kTile;
  using Pk = Pack16<T>;
  constexpr int PN = Pk::N;         // rows per pack
  constexpr int PL = kTile / PN;    // threads per column
  constexpr int PC = 256 / PL;      // columns per pass of the workgroup
  constexpr int NP = KC / PC;       // packs per thread and operand
  static_assert(!packed || KC % PC == 0, "pack passes");
  const int pi = PN * (tid % PL), pt = tid / PL;
  const int rap = r0 + pi, rbp = c0 + pi;
  const int64_t rapS = rap < rlim ? rap : r0, rbpS = rbp < m ? rbp : c0;
  Pk ap[packed ? NP : 1], bp[packed ? NP : 1];
  T dp[packed ? NP : 1];
  // entry-wise path (KC = 64): thread (row i = lane, columns t = w + 4q)
  const int ra = r0 + lane, rb = c0 + lane;
  const int64_t ras = ra < rlim ? ra : r0, rbs = rb < m ? rb : c0;
  T av[packed ? 1 : 16], bv[packed ? 1 : 16], dv[packed ? 1 : 16];
  auto load = [&](int q) {
    const int k0 = q * KC;
    const int nbq = min(KC, nc - k0);
    if constexpr (packed) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int t = min(pt + PC * j, nbq - 1);
        ap[j] = load_pack16(P + rapS + (k0 + t) * ld);
        bp[j] = load_pack16(P + rbpS + (k0 + t) * ld);
        dp[j] = a.D[f + k0 + t];
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int t = min(w + 4 * u, nbq - 1);
        av[u] = M::ld(P + ras + (k0 + t) * ld);
        bv[u] = M::ld(P + rbs + (k0 + t) * ld);
        dv[u] = M::ld(a.D + f + k0 + t);
      }
    }
  };
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  V acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[x][y][v] = T(0);
  // chunks of the steps kfirst .. klast that hold columns of the front
  constexpr int CPS = kPanelNB / KC;
  const int qfirst = kfirst * CPS;
  const int qlast = min(klast * CPS + CPS - 1, (nc - 1) / KC);
  load(qfirst);
  for (int q = qfirst; q <= qlast; ++q) {
    const int nbq = min(KC, nc - q * KC);
    if constexpr (packed) {
      // PN consecutive rows stay consecutive under the swizzle (it flips
      // bits 4 and 5 of the row index only): one 16-byte LDS store each
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int t = pt + PC * j;
        Pk x, y;
#pragma unroll
        for (int e = 0; e < PN; ++e) {
          x.e[e] = (rap + e < rlim && t < nbq) ? ap[j].e[e] : T(0);
          y.e[e] = (rbp + e < m && t < nbq) ? bp[j].e[e] * dp[j] : T(0);
        }
        *reinterpret_cast<Pk*>(&L0[at(t, pi)]) = x;
        *reinterpret_cast<Pk*>(&L1[at(t, pi)]) = y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int t = w + 4 * u;
        L0[at(t, lane)] = (ra < rlim && t < nbq) ? av[u] : T(0);
        L1[at(t, lane)] = (rb < m && t < nbq) ? bv[u] * dv[u] : T(0);
      }
    }
    __syncthreads();
    if (q < qlast) load(q + 1);
    for (int t0 = 0; t0 < nbq; t0 += 4) {
      const int t = t0 + lk;  // rows t >= nbq of the operand tiles are zero
      const T a0 = L0[at(t, wr + li)], a1 = L0[at(t, wr + 16 + li)];
      const T b0 = L1[at(t, wc + li)], b1 = L1[at(t, wc + 16 + li)];
      acc[0][0] = Mfma16<T>::op(a0, b0, acc[0][0]);
      acc[0][1] = Mfma16<T>::op(a0, b1, acc[0][1]);
      acc[1][0] = Mfma16<T>::op(a1, b0, acc[1][0]);
      acc[1][1] = Mfma16<T>::op(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  // the tile goes through LDS (Ls is free after the last barrier, viewed as
  // St[column][row] with row stride kTile + 1, so that a wave's reads of one
  // column are conflict-free) so that the read-modify-write of the target is
  // coalesced: element (x, y, v) is row wr + 16x + row(lk, v), column
  // wc + 16y + li of the tile
  T(*St)[kTile + 1] = reinterpret_cast<T(*)[kTile + 1]>(Ls);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) St[wc + 16 * y + li][wr + 16 * x + Mfma16<T>::row(lk, v)] = acc[x][y][v];
  __syncthreads();
  // thread (row lane, columns w + 4q): every load first, then the stores
  const int r = r0 + lane;
  T* dst[16];
  T cur[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = c0 + w + 4 * q;
    const bool ok = c < m && r < rlim && r >= c && (c < nc || (r < m && !panel_only));
    T* p = !ok ? Pw : c < nc ? Pw + r + c * ld : U + (r - nc) + static_cast<int64_t>(c - nc) * nr;
    cur[q] = M::ld(p);
    dst[q] = ok ? p : nullptr;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q)
    if (dst[q] != nullptr) M::st(dst[q], cur[q] - St[w + 4 * q][lane]);
}

template <class T, int R>
__global__ void __launch_bounds__(256) k_update_tiles(FactorArgs<T> a, const Task4* __restrict__ tasks) {
  __shared__ __align__(16) T Ls[upd_lds<PlainMem>()];
  // Workgroups are dealt round-robin over the XCDs (blockIdx % kXcds); each
  // XCD takes a contiguous run of the task list instead, so tiles that share
  // operand rows and columns meet in the same L2.
  upd_body<T, R, PlainMem>(a, tasks[xcd_task(blockIdx.x, gridDim.x)], Ls);
}

// Trailing update tile for a factor held in a 16-bit format (R = 1 FP16,
// R = 2 BF16): the task layout and the FP32 read-modify-write of
// k_update_tiles, with the operand tiles in that format on the 16-bit matrix
// cores (v_mfma_f32_16x16x32, FP32 accumulation). L entries are already
// rounded to the format as stored; the column operand L_c D is rounded to it
// as it is formed (the FP32 path keeps that product in FP32). Operand tiles
// are row-major with k contiguous: lane l of an MFMA reads the 8 k-values
// 8 (l / 16) .. + 7 of row l % 16 as one 16-byte load, and the row stride of
// 72 entries (36 banks) puts the 16 rows of a read in distinct bank groups.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
template <int R>
struct Mfma16x32;
template <>
struct Mfma16x32<1> {
  using E = _Float16;
  using V8 = f16x8;
  static __device__ __forceinline__ f32x4 op(V8 a, V8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma16x32<2> {
  using E = __bf16;
  using V8 = bf16x8;
  static __device__ __forceinline__ f32x4 op(V8 a, V8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

constexpr int kLs16 = kPanelNB + 8;  // row stride (entries) of the 16-bit operand tiles

// Ls: 2 x kTile x kLs16 entries of LDS (16-byte aligned); Ls[0][i * LS + t] =
// L[r0 + i, k0 + t], Ls[1]: L[c0 + j, k0 + t] D.
template <int R, class Mem>
__device__ __forceinline__ void upd16_body(const FactorArgs<float>& a, const Task4 tk,
                                           typename Mfma16x32<R>::E (*Ls)[kTile * kLs16]) {
  using M = Mfma16x32<R>;
  using E = typename M::E;
  using V8 = typename M::V8;
  constexpr int LS = kLs16;
  static_assert(kTile * (kTile + 1) * sizeof(float) <= 2 * kTile * LS * sizeof(E), "staging tile exceeds the operands");
  const int s = tk.a, r0 = tk.b, c0 = tk.c;
  const int klast = tk.d & ((1 << kUpdStepsShift) - 1);
  const int kfirst = klast - ((tk.d >> kUpdStepsShift) & ((1 << (kUpdPanelBit - kUpdStepsShift)) - 1)) + 1;
  const bool panel_only = (tk.d >> kUpdPanelBit) & 1;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int rlim = r0 >= m ? m + nc : m;
  const float* P = a.panel + a.panel_off[s];
  float* Pw = a.panel + a.panel_off[s];
  float* U = nr > 0 ? a.upool + a.upd_off[s] : Pw;
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ra = r0 + lane, rb = c0 + lane;
  const int64_t ras = ra < rlim ? ra : r0, rbs = rb < m ? rb : c0;
  float av[16], bv[16], dv[16];
  auto load = [&](int kk) {
    const int k0 = kk * kPanelNB;
    const int nbk = min(kPanelNB, nc - k0);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int t = min(w + 4 * q, nbk - 1);
      av[q] = Mem::ld(P + ras + (k0 + t) * ld);
      bv[q] = Mem::ld(P + rbs + (k0 + t) * ld);
      dv[q] = Mem::ld(a.D + f + k0 + t);
    }
  };
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[x][y][v] = 0.f;
  load(kfirst);
  for (int kk = kfirst; kk <= klast; ++kk) {
    const int nbk = min(kPanelNB, nc - kk * kPanelNB);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int t = w + 4 * q;
      Ls[0][lane * LS + t] = static_cast<E>((ra < rlim && t < nbk) ? av[q] : 0.f);
      Ls[1][lane * LS + t] = static_cast<E>((rb < m && t < nbk) ? bv[q] * dv[q] : 0.f);
    }
    __syncthreads();
    if (kk < klast) load(kk + 1);
#pragma unroll
    for (int t0 = 0; t0 < kPanelNB; t0 += 32) {  // entries t >= nbk of the operand rows are zero
      const V8 a0 = *reinterpret_cast<const V8*>(&Ls[0][(wr + li) * LS + t0 + 8 * lk]);
      const V8 a1 = *reinterpret_cast<const V8*>(&Ls[0][(wr + 16 + li) * LS + t0 + 8 * lk]);
      const V8 b0 = *reinterpret_cast<const V8*>(&Ls[1][(wc + li) * LS + t0 + 8 * lk]);
      const V8 b1 = *reinterpret_cast<const V8*>(&Ls[1][(wc + 16 + li) * LS + t0 + 8 * lk]);
      acc[0][0] = M::op(a0, b0, acc[0][0]);
      acc[0][1] = M::op(a0, b1, acc[0][1]);
      acc[1][0] = M::op(a1, b0, acc[1][0]);
      acc[1][1] = M::op(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  // staging and read-modify-write as in k_update_tiles (output layout of the
  // 16x16x32 products: row 4 (l / 16) + v, column l % 16)
  float(*St)[kTile + 1] = reinterpret_cast<float(*)[kTile + 1]>(&Ls[0][0]);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) St[wc + 16 * y + li][wr + 16 * x + 4 * lk + v] = acc[x][y][v];
  __syncthreads();
  const int r = r0 + lane;
  float* dst[16];
  float cur[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = c0 + w + 4 * q;
    const bool ok = c < m && r < rlim && r >= c && (c < nc || (r < m && !panel_only));
    float* p = !ok ? Pw : c < nc ? Pw + r + c * ld : U + (r - nc) + static_cast<int64_t>(c - nc) * nr;
    cur[q] = Mem::ld(p);
    dst[q] = ok ? p : nullptr;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q)
    if (dst[q] != nullptr) Mem::st(dst[q], cur[q] - St[w + 4 * q][lane]);
}

template <int R>
__global__ void __launch_bounds__(256) k_update_tiles16(FactorArgs<float> a, const Task4* __restrict__ tasks) {
  __shared__ __align__(16) typename Mfma16x32<R>::E Ls[2][kTile * kLs16];
  upd16_body<R, PlainMem>(a, tasks[xcd_task(blockIdx.x, gridDim.x)], Ls);
}

// ---------------------------------------------------------------- dependency-driven factorization
// Diagonal block of panel step k of big front s, once, by 256 threads
// (thread (row lane, wave w) holds columns 16 w .. 16 w + 15): the
// elimination of k_panel_step's owning task, step for step (4 columns per
// barrier, the wave holding them factors the 4 x 4 pivot block and publishes
// the multipliers; every entry takes the same operations in the same order,
// so the results are those of the per-level launches bit for bit). Writes
// L_kk (transposed, in the block's upper triangle), D, D^{-1} and
// Z_kk = L_kk^{-T} D_k^{-1} in the block's identity rows, with its zeros as
// the per-level launches leave them (the product I Z_kk: +0).
// lds: 6 * 64 * 4 + 64 * 65 entries.
template <class T, int R>
__device__ __forceinline__ void diag_body(const FactorArgs<T>& a, int s, int k, T* __restrict__ lds) {
  constexpr int NB = kPanelNB;
  constexpr int CW = 4;       // columns per elimination step
  constexpr int QW = NB / 4;  // columns per wave
  T(*mrs)[NB][CW] = reinterpret_cast<T(*)[NB][CW]>(lds);
  T(*gms)[NB][CW] = mrs + 2;
  T(*dms)[NB][CW] = mrs + 4;
  T(*Xs)[NB + 1] = reinterpret_cast<T(*)[NB + 1]>(lds + 6 * NB * CW);
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int64_t ld = panel_ld(nc, a.nrow[s]);
  const int k0 = k * NB;
  const int nbk = min(NB, nc - k0);
  T* P = a.panel + a.panel_off[s];
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int c0w = QW * w;
  const int lsafe = min(lane, nbk - 1);
  T fa[QW], ga[QW];
#pragma unroll
  for (int u = 0; u < QW; ++u) {
    const int c = c0w + u;
    const int cs = min(c, nbk - 1);
    const T v = ld_dev(P + (k0 + lsafe) + (k0 + cs) * ld);
    const bool in = lane < nbk && c < nbk && lane >= c;
    fa[u] = in ? v : T(0);
    ga[u] = (lane == c && c < nbk) ? T(1) : T(0);
  }
  for (int wo = 0; wo < 4; ++wo) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int J = 4 * wo + jj;
      const int b = J & 1;
      const int j0 = CW * J;
      if (w == wo) {
        bool valid[CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) valid[t] = j0 + t < nbk;
        // pivot block L_B D_B L_B^T; dl[t][q] = d_q l_tq, dinv = 1 / d
        T d[CW], dinv[CW], l[CW][CW], dl[CW][CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          T dt = lane_value(fa[4 * jj + t], j0 + t);
#pragma unroll
          for (int q = 0; q < t; ++q) dt -= l[t][q] * dl[t][q];
          d[t] = valid[t] ? dt : T(1);
          dinv[t] = recip_nr(d[t]);
#pragma unroll
          for (int u = t + 1; u < CW; ++u) {
            T v = lane_value(fa[4 * jj + t], min(j0 + u, NB - 1));
#pragma unroll
            for (int q = 0; q < t; ++q) v -= l[u][q] * dl[t][q];
            v *= dinv[t];
            l[u][t] = (valid[u] && valid[t]) ? v : T(0);
            dl[u][t] = d[t] * l[u][t];
          }
        }
        // multipliers of this lane's row and identity row
        T mr[CW], gm[CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          T v = fa[4 * jj + t];
          T g = ga[4 * jj + t];
#pragma unroll
          for (int q = 0; q < t; ++q) {
            v -= mr[q] * dl[t][q];
            g -= gm[q] * dl[t][q];
          }
          mr[t] = v * dinv[t];
          gm[t] = g * dinv[t];
        }
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          mrs[b][lane][t] = mr[t];
          gms[b][lane][t] = gm[t];
          dms[b][lane][t] = d[t] * mr[t];
        }
        // columns of L_kk (and D on the diagonal) into LDS
        const int s0 = lane - j0;
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          T lv = T(0);
#pragma unroll
          for (int u = t + 1; u < CW; ++u) lv = s0 == u ? l[u][t] : lv;
          Xs[j0 + t][lane] = s0 < t ? T(0) : s0 == t ? d[t] : s0 < CW ? lv : mr[t];
        }
#pragma unroll
        for (int t = 0; t < CW; ++t)
          if (valid[t]) ga[4 * jj + t] = gm[t];
      }
      __syncthreads();
      // rank-CW update of the columns right of the step's block
      if (w >= wo) {
        T mr[CW], gm[CW];
#pragma unroll
        for (int t = 0; t < CW; ++t) {
          mr[t] = mrs[b][lane][t];
          gm[t] = gms[b][lane][t];
        }
#pragma unroll
        for (int u = 0; u < QW; ++u) {
          if (w == wo && u < 4 * jj + 4) continue;
          const int c = c0w + u;
          T dm[CW];
#pragma unroll
          for (int t = 0; t < CW; ++t) dm[t] = dms[b][c][t];
          T fu = T(0), gu = T(0);
#pragma unroll
          for (int t = 0; t < CW; ++t) {
            fu += mr[t] * dm[t];
            gu += gm[t] * dm[t];
          }
          const bool on = c < nbk;
          fa[u] = (on && lane >= c) ? fa[u] - fu : fa[u];
          ga[u] = on ? ga[u] - gu : ga[u];
          // four columns at a time: the LDS reads of all sixteen in flight
          // at once would take 64 more registers
          if ((u & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
  __syncthreads();
  for (int rr = w; rr < nbk; rr += 4)
    if (lane < rr) st_dev(P + (k0 + lane) + (k0 + rr) * ld, rnd<R>(Xs[lane][rr]));
  if (tid < nbk) {
    const T d = Xs[tid][tid];
    if (d == T(0)) flag_zero_pivot(a.err);
    st_dev(a.D + f + k0 + tid, d);
    st_dev(a.Dinv + f + k0 + tid, T(1) / d);
  }
#pragma unroll
  for (int u = 0; u < QW; ++u) {
    const int c = c0w + u;
    if (lane < nbk && c < nbk) st_dev(P + (m + k0 + lane) + (k0 + c) * ld, ga[u] + T(0));
  }
}

// Rows [r0, r1) (at most 64) of panel step k of big front s: L = X Z_kk on the
// matrix cores, X = the rows' current entries in the step's columns (rows
// below the diagonal block, or identity rows of earlier blocks), Z_kk from
// the block's identity rows. Every output entry takes the MFMA sequence of
// k_panel_step (products over t = 0, 4, 8, ... in order). Ls: 2 x kPanelNB x
// kTile entries of LDS.
template <class T, int R>
__device__ __forceinline__ void pan_body(const FactorArgs<T>& a, int s, int r0, int r1, int k,
                                         T (*Ls)[kPanelNB * kTile]) {
  using V = typename Mfma16<T>::V;
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int m = nc + a.nrow[s];
  const int64_t ld = panel_ld(nc, a.nrow[s]);
  const int k0 = k * kPanelNB;
  const int nbk = min(kPanelNB, nc - k0);
  T* P = a.panel + a.panel_off[s];
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rows = r1 - r0;
  const int r = r0 + lane;
  const int64_t rs = lane < rows ? r : r0;
  const int zl = min(lane, nbk - 1);
  // thread (lane, t = w + 4 q): X[r0 + lane][t] and Z_kk[lane][t]
  T xv[16], zv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int ts = min(w + 4 * q, nbk - 1);
    xv[q] = ld_dev(P + rs + (k0 + ts) * ld);
    zv[q] = ld_dev(P + (m + k0 + zl) + (k0 + ts) * ld);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int t = w + 4 * q;
    Ls[0][op_at(t, lane)] = (lane < rows && t < nbk) ? xv[q] : T(0);   // X^T: [t][row]
    Ls[1][op_at(lane, t)] = (lane < nbk && t < nbk) ? zv[q] : T(0);    // Z: [row][column]
  }
  __syncthreads();
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  V acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[x][y][v] = T(0);
  for (int t0 = 0; t0 < nbk; t0 += 4) {
    const int t = t0 + lk;
    const T a0 = Ls[0][op_at(t, wr + li)], a1 = Ls[0][op_at(t, wr + 16 + li)];
    const T b0 = Ls[1][op_at(t, wc + li)], b1 = Ls[1][op_at(t, wc + 16 + li)];
    acc[0][0] = Mfma16<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma16<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma16<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma16<T>::op(a1, b1, acc[1][1]);
  }
  __syncthreads();
  T(*St)[kTile + 1] = reinterpret_cast<T(*)[kTile + 1]>(&Ls[0][0]);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) St[wc + 16 * y + li][wr + 16 * x + Mfma16<T>::row(lk, v)] = acc[x][y][v];
  __syncthreads();
  if (lane >= rows) return;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = w + 4 * q;
    const T v = St[c][lane];
    if (c < nbk) st_dev(P + r + (k0 + c) * ld, r < m ? rnd<R>(v) : v);
  }
}

// A bounded wait of the dependency-driven factorization (as spin_until);
// once any wait of the launch has expired (bit 2 of *err), the others end
// too, so a failing launch drains in microseconds instead of one bound per
// task. Task indices come from the plan, never from computed values, so
// such a launch stays within its buffers.
__device__ __forceinline__ void dag_wait(const int* c, int target, int* err) {
  int it = 0;
  while (ld_dev(c) < target) {
    __builtin_amdgcn_s_sleep(1);
    ++it;
    if ((it & 255) == 0 && (ld_dev(err) & 2)) break;
    if (it > kSpinLimit) {
      atomicOr(err, 2);
      break;
    }
  }
}

// The levels >= the plan's fdag_level in one persistent launch: workgroups
// take tasks from the dispensers in list order (topological), wait on the
// counters the task names, run it and count it (see sync.hpp for why the
// launch always progresses). Every value another task of the launch wrote is
// read and written with device-coherent accesses.
template <class T, int R>
__global__ void __launch_bounds__(256) k_factor_dag(FactorArgs<T> a, FactorDagDev g) {
  __shared__ __align__(16) T smem[2 * kPanelNB * kTile];
  __shared__ int tsh;
  static_assert(6 * kPanelNB * 4 + kPanelNB * (kPanelNB + 1) <= 2 * kPanelNB * kTile, "diagonal-block LDS");
  static_assert(kAsmTile * (kAsmTile + 1) <= 2 * kPanelNB * kTile, "assembly LDS");
  static_assert(2 * kTile * kLs16 * 2 <= 2 * kPanelNB * kTile * sizeof(T), "16-bit operand LDS");
  int* cnt = g.counters;
  int* lines = cnt + g.cnt_disp;
  const FactorDagIndex X{g.ns, g.nsteps};
  Dispenser ds = arrive(lines, &tsh);
  for (;;) {
    const int t = next_task(lines, ds, g.ntask, &tsh);
    if (t >= g.ntask) break;
    const FTask tk = g.task[t];
    const int kind = tk.sk >> kFKindShift;
    const int s = tk.sk & ((1 << kFKindShift) - 1);
    unsigned long long t_draw = 0;
    if (threadIdx.x == 0) {
      if (g.trace) t_draw = wall_clock64();
      dag_wait(cnt + tk.wa, tk.ta, a.err);
      if (tk.wb >= 0) dag_wait(cnt + tk.wb, tk.tb, a.err);
      if (tk.wc >= 0) dag_wait(cnt + tk.wc, tk.tc, a.err);
      if (g.trace) {
        g.trace[4 * static_cast<int64_t>(t)] = static_cast<unsigned long long>(tk.sk);
        g.trace[4 * static_cast<int64_t>(t) + 1] = t_draw;
        g.trace[4 * static_cast<int64_t>(t) + 2] = wall_clock64();
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    __syncthreads();
    const int* ms = g.meta + kFMetaStride * s;
    if (kind == kFAsm) {
      asm_body<T, R, CoherentMem, 8>(a, g.asm_task[tk.x], g.asm_child_ptr[tk.x], g.asm_child_ptr[tk.x + 1],
                                     g.asm_child, reinterpret_cast<T(*)[kAsmTile + 1]>(smem));
      block_publish(cnt + X.asm_done(s));
    } else if (kind == kFDiag) {
      diag_body<T, R>(a, s, tk.x, smem);
      if (block_publish(cnt + X.diag_done(ms[kFMetaStep] + tk.x)) >= 0) add_dev(cnt + X.pan_done(ms[kFMetaStep] + tk.x), 1);
    } else if (kind == kFPan) {
      pan_body<T, R>(a, s, tk.x, tk.y, tk.z, reinterpret_cast<T(*)[kPanelNB * kTile]>(smem));
      block_publish(cnt + X.pan_done(ms[kFMetaStep] + tk.z));
    } else {
      const Task4 ut{s, tk.x, tk.y, tk.z};
      if constexpr (std::is_same<T, float>::value && R != 0)
        upd16_body<R, CoherentMem>(a, ut, reinterpret_cast<typename Mfma16x32<R>::E(*)[kTile * kLs16]>(smem));
      else
        upd_body<T, R, CoherentMem>(a, ut, reinterpret_cast<T*>(smem));
      // counted on its column block (tiles that start on one), then towards
      // the front's completion, which the last tile reports to the parent
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (threadIdx.x == 0) {
        if (tk.y % kTile == 0) add_dev(cnt + X.col_done(ms[kFMetaCb] + tk.y / kTile), 1);
        if (add_dev(cnt + X.upd_done(s), 1) == ms[kFMetaNUpd] - 1 && ms[kFMetaParent] >= 0)
          add_dev(cnt + X.child_done(ms[kFMetaParent]), 1);
      }
    }
    if (g.trace && threadIdx.x == 0) g.trace[4 * static_cast<int64_t>(t) + 3] = wall_clock64() << 16 | blockIdx.x;
  }
}

__global__ void k_fill_zero(int n, int* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Workgroups of a kernel resident on the whole device at once.
template <class K>
int resident_workgroups(K kernel, int threads) {
  int dev = 0, cus = 0, per_cu = 0;
  MPIR_HIP(hipGetDevice(&dev));
  MPIR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  MPIR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  return std::max(1, per_cu) * std::max(1, cus);
}

// Count of D[i] > 0 (the inertia QDLDL_factor returns) and of i with D[i]
// or Dinv[i] not finite: one atomic per counter and workgroup (256 threads),
// few workgroups, so that the adds to one counter do not queue up. An
// overflow or NaN anywhere in a column of L reaches the pivot of the row it
// sits in (D_i -= L_ij^2 D_j), so the pivots tell whether the factor is finite.
template <class T>
__global__ void __launch_bounds__(256) k_count_positive(int n, const T* __restrict__ D, const T* __restrict__ Dinv,
                                                        int* __restrict__ out, int* __restrict__ nonfinite) {
  __shared__ int part[8];
  int cnt = 0, bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T d = D[i];
    cnt += d > T(0) ? 1 : 0;
    bad += (isfinite(d) && isfinite(Dinv[i])) ? 0 : 1;
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_down(cnt, off);
    bad += __shfl_down(bad, off);
  }
  if ((threadIdx.x & 63) == 0) {
    part[threadIdx.x >> 6] = cnt;
    part[4 + (threadIdx.x >> 6)] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int c = part[0] + part[1] + part[2] + part[3];
    const int b = part[4] + part[5] + part[6] + part[7];
    if (c) atomicAdd(out, c);
    if (b) atomicAdd(nonfinite, b);
  }
}

inline unsigned blocks_for(int64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

// the trailing-update launch: 16-bit operands on the 16-bit matrix cores for
// a factor held in FP16 / BF16, the factor's own precision otherwise
template <class T, int R>
void launch_update(const FactorArgs<T>& a, const Task4* tasks, int count, hipStream_t st) {
  if constexpr (std::is_same<T, float>::value && R != 0)
    hipLaunchKernelGGL((k_update_tiles16<R>), dim3(count), dim3(256), 0, st, a, tasks);
  else
    hipLaunchKernelGGL((k_update_tiles<T, R>), dim3(count), dim3(256), 0, st, a, tasks);
}

template <class T, int R>
void factor_levels(const FactorArgs<T>& a, const DeviceFactorPlan& dp, const double* ax_scaled, int64_t nnz,
                   hipStream_t st) {
  MPIR_HIP(hipMemsetAsync(a.panel, 0, sizeof(T) * dp.panel_entries, st));
  MPIR_HIP(hipMemsetAsync(a.err, 0, sizeof(int), st));
  if (nnz > 0)
    hipLaunchKernelGGL((k_scatter_values<T, R>), dim3(blocks_for(nnz, 256)), dim3(256), 0, st, nnz, ax_scaled,
                       dp.val_dst, a.panel);
  // levels below the dependency-driven launch (all of them without it):
  // one launch per level and kind, and per step of the big fronts
  const bool dag = dp.use_dag && dp.dag_pool && dp.dag_level < dp.nlevels && dp.dag.ntask > 0;
  const int lsplit = dag ? dp.dag_level : dp.nlevels;
  for (int l = 0; l < lsplit; ++l) {
    const int s0 = dp.h_small_ptr[l], s1 = dp.h_small_ptr[l + 1];
    if (s1 > s0)
      hipLaunchKernelGGL((k_small_fronts<T, R>), dim3(s1 - s0), dim3(64), 0, st, a, dp.small_list + s0);
    const int a0 = dp.h_asm_ptr[l], a1 = dp.h_asm_ptr[l + 1];
    if (a1 > a0)
      hipLaunchKernelGGL((k_assemble<T, R>), dim3(a1 - a0), dim3(256), 0, st, a, dp.asm_task + a0,
                         dp.asm_child_ptr + a0, dp.asm_child);
    const int st0 = dp.h_level_step_ptr[l], st1 = dp.h_level_step_ptr[l + 1];
    const bool split_level = a.zbuf != nullptr && dp.h_big_ptr[l + 1] - dp.h_big_ptr[l] <= dp.zbuf_fronts;
    for (int q = st0; q < st1; ++q) {
      const int p0 = dp.h_pan_ptr[q], p1 = dp.h_pan_ptr[q + 1];
      if (split_level && dp.pan_split_min > 0 && p1 - p0 >= dp.pan_split_min) {
        // the fronts' diagonal blocks, then every other task with Z_kk read
        // back (each from the step's own list)
        const int o0 = dp.h_pan_own_ptr[q], o1 = dp.h_pan_own_ptr[q + 1];
        const int e0 = dp.h_pan_rest_ptr[q], e1 = dp.h_pan_rest_ptr[q + 1];
        if (o1 > o0)
          hipLaunchKernelGGL((k_panel_step<T, R, true, 1>), dim3(o1 - o0), dim3(kPanelThreads), 0, st, a,
                             dp.pan_own + o0);
        if (e1 > e0)
          hipLaunchKernelGGL((k_panel_step<T, R, true, 2>), dim3(e1 - e0), dim3(kPanelThreads), 0, st, a,
                             dp.pan_rest + e0);
      } else if (p1 > p0) {
        hipLaunchKernelGGL((k_panel_step<T, R, true>), dim3(p1 - p0), dim3(kPanelThreads), 0, st, a,
                           dp.pan_task + p0);
      }
      const int u0 = dp.h_upd_ptr[q], u1 = dp.h_upd_ptr[q + 1];
      if (u1 > u0) launch_update<T, R>(a, dp.upd_task + u0, u1 - u0, st);
    }
  }
  if (dag) {
    const FactorDagDev& g = dp.dag;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(g.ncounters, 256)), dim3(256), 0, st, g.ncounters, g.counters);
    // at least kDispensers workgroups, so that every class is covered
    static const int grid = resident_workgroups(k_factor_dag<T, R>, 256);
    hipLaunchKernelGGL((k_factor_dag<T, R>), dim3(std::max(kDispensers, std::min(g.ntask, grid))), dim3(256), 0, st, a,
                       g);
  }
}

// W = Z D: column c of the inverse rows of every big front times D_c, which
// turns Z = L_11^{-T} D^{-1} into the unit upper triangular L_11^{-T}.
// Tasks (s, column block j, row chunk h) cover the rows at or above the
// block's last column.
template <class T>
__global__ void __launch_bounds__(256) k_inverse_to_unit(FactorArgs<T> a, const int32_t* __restrict__ tasks) {
  const int s = tasks[3 * blockIdx.x], j = tasks[3 * blockIdx.x + 1], h = tasks[3 * blockIdx.x + 2];
  const int f = a.first[s];
  const int nc = a.first[s + 1] - f;
  const int nr = a.nrow[s];
  const int m = nc + nr;
  const int64_t ld = panel_ld(nc, nr);
  const int c0 = j * kPanelNB;
  const int bw = min(kPanelNB, nc - c0);
  const int q0 = h * kXcRows;
  const int len = min(kXcRows, c0 + bw - q0);
  T* W = a.panel + a.panel_off[s] + m + q0;
  for (int e = threadIdx.x; e < len * bw; e += blockDim.x) {
    const int q = e % len, c = c0 + e / len;
    W[q + c * ld] *= a.D[f + c];
  }
}

template <class T>
void factor_numeric(const FactorArgs<T>& a, const DeviceFactorPlan& dp, const double* ax_scaled, int64_t nnz,
                    int rounding, hipStream_t st) {
  if constexpr (std::is_same<T, float>::value) {
    if (rounding == 1) factor_levels<T, 1>(a, dp, ax_scaled, nnz, st);
    else if (rounding == 2) factor_levels<T, 2>(a, dp, ax_scaled, nnz, st);
    else factor_levels<T, 0>(a, dp, ax_scaled, nnz, st);
  } else {
    factor_levels<T, 0>(a, dp, ax_scaled, nnz, st);
  }
  if (dp.n_inv_task > 0)
    hipLaunchKernelGGL(k_inverse_to_unit<T>, dim3(dp.n_inv_task), dim3(256), 0, st, a, dp.inv_task);
  MPIR_HIP(hipGetLastError());
}

template <class T>
void invert_diag_blocks(const FactorArgs<T>& a, const DeviceFactorPlan& dp, hipStream_t st) {
  for (int l = 0; l < dp.nlevels; ++l) {
    const int st0 = dp.h_level_step_ptr[l], st1 = dp.h_level_step_ptr[l + 1];
    for (int q = st0; q < st1; ++q) {
      const int p0 = dp.h_pan_aug_ptr[q], p1 = dp.h_pan_ptr[q + 1];
      if (p1 > p0)
        hipLaunchKernelGGL((k_panel_step<T, 0, false>), dim3(p1 - p0), dim3(kPanelThreads), 0, st, a,
                           dp.pan_task + p0);
      const int u0 = dp.h_upd_aug_ptr[q], u1 = dp.h_upd_ptr[q + 1];
      if (u1 > u0) hipLaunchKernelGGL((k_update_tiles<T, 0>), dim3(u1 - u0), dim3(256), 0, st, a, dp.upd_task + u0);
    }
  }
  if (dp.n_inv_task > 0)
    hipLaunchKernelGGL(k_inverse_to_unit<T>, dim3(dp.n_inv_task), dim3(256), 0, st, a, dp.inv_task);
  MPIR_HIP(hipGetLastError());
}

// The factor in a 16-bit format: dst[q] = round(src[q]) over the whole
// panel pool (FP16 or BF16, round to nearest even).
template <class H>
__global__ void k_round_panel(int64_t n, const float* __restrict__ src, H* __restrict__ dst) {
  const int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (q < n) {
    if constexpr (std::is_same<H, __half>::value) dst[q] = __float2half_rn(src[q]);
    else dst[q] = __float2bfloat16(src[q]);
  }
}

// counts entries of the rounded panel that are not finite (out of range)
template <class H>
__global__ void k_count_nonfinite(int64_t n, const H* __restrict__ v, int* __restrict__ out) {
  int cnt = 0;
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < n;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x)
    cnt += isfinite(static_cast<float>(v[q])) ? 0 : 1;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, cnt);
}

// adds the count of non-finite rounded entries to *nonfinite
template <class H>
void round_panel(int64_t n, const float* src, H* dst, int* nonfinite, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_round_panel<H>, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, src, dst);
  hipLaunchKernelGGL(k_count_nonfinite<H>, dim3(std::min<unsigned>(2048, blocks_for(n, 256))), dim3(256), 0, st, n,
                     dst, nonfinite);
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

// *out = #{D_i > 0}, *nonfinite = #{D_i or Dinv_i not finite} (both reset)
template <class T>
void count_positive(int n, const T* D, const T* Dinv, int* out, int* nonfinite, hipStream_t st) {
  MPIR_HIP(hipMemsetAsync(out, 0, sizeof(int), st));
  MPIR_HIP(hipMemsetAsync(nonfinite, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_count_positive<T>, dim3(std::min<unsigned>(256, std::max(1u, blocks_for(n, 256)))), dim3(256), 0,
                     st, n, D, Dinv, out, nonfinite);
}

template void factor_numeric<double>(const FactorArgs<double>&, const DeviceFactorPlan&, const double*, int64_t, int,
                                     hipStream_t);
template void factor_numeric<float>(const FactorArgs<float>&, const DeviceFactorPlan&, const double*, int64_t, int,
                                    hipStream_t);
template void invert_diag_blocks<double>(const FactorArgs<double>&, const DeviceFactorPlan&, hipStream_t);
template void invert_diag_blocks<float>(const FactorArgs<float>&, const DeviceFactorPlan&, hipStream_t);
template void round_panel<__half>(int64_t, const float*, __half*, int*, hipStream_t);
template void round_panel<__hip_bfloat16>(int64_t, const float*, __hip_bfloat16*, int*, hipStream_t);
template void reciprocal<double>(int, const double*, double*, hipStream_t);
template void reciprocal<float>(int, const float*, float*, hipStream_t);
template void count_positive<double>(int, const double*, const double*, int*, int*, hipStream_t);
template void count_positive<float>(int, const float*, const float*, int*, int*, hipStream_t);

}  // namespace mpir
