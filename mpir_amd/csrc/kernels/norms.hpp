// Device building blocks of the vector norms of GMRES-IR.
//
// Dnrm2 (include/MPIR/ops.hpp:231-255 of yex33/MPIR) is LAPACK's scaled sum
// of squares: the value is scale * sqrt(ssq) with every term divided by the
// largest magnitude seen so far, so it neither overflows for entries near
// 1e308 nor underflows for entries near 1e-308. Here each thread keeps such a
// (scale, ssq) pair over its entries and the pairs are merged up a reduction
// tree; a NaN entry makes the norm NaN, as in the reference.
//
// InfNrm (ops.hpp:324-329) is max |x_i|. The reference's std::max_element
// returns NaN only when x[0] is NaN and skips a NaN elsewhere; the device
// maximum returns NaN when any entry is NaN, so a non-finite vector is never
// mistaken for a finite one.
#pragma once

#include <hip/hip_runtime.h>

namespace mpir {

// max(a, b) that returns NaN when either operand is NaN
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

struct Ssq {
  double scale;  // largest |x_i| so far (0 before the first nonzero)
  double ssq;    // sum of (x_i / scale)^2, at least 1 once scale > 0
};

__device__ __forceinline__ Ssq ssq_init() { return Ssq{0.0, 1.0}; }

// one entry, ops.hpp:243-253
__device__ __forceinline__ Ssq ssq_add(Ssq a, double x) {
  if (x != 0.0) {
    const double ax = fabs(x);
    if (a.scale < ax) {
      const double r = a.scale / ax;
      a.ssq = 1.0 + a.ssq * (r * r);
      a.scale = ax;
    } else {
      const double r = ax / a.scale;
      a.ssq = a.ssq + r * r;
    }
  }
  return a;
}

// two partial sums: rescale the one with the smaller scale
__device__ __forceinline__ Ssq ssq_merge(Ssq a, Ssq b) {
  if (a.scale != a.scale || a.ssq != a.ssq) return a;
  if (b.scale != b.scale || b.ssq != b.ssq) return b;
  if (a.scale < b.scale) {
    const Ssq t = a;
    a = b;
    b = t;
  }
  if (b.scale == 0.0) return a;
  const double r = b.scale / a.scale;
  a.ssq = a.ssq + b.ssq * (r * r);
  return a;
}

__device__ __forceinline__ double ssq_value(Ssq a) { return a.scale == 0.0 ? 0.0 : a.scale * sqrt(a.ssq); }

// wave then block merge of per-thread pairs (blockDim.x a multiple of 64,
// sh holds 2 * blockDim.x / 64 doubles); the result is valid in thread 0
__device__ __forceinline__ Ssq block_ssq(Ssq v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) {
    Ssq o;
    o.scale = __shfl_down(v.scale, off);
    o.ssq = __shfl_down(v.ssq, off);
    v = ssq_merge(v, o);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sh[2 * w] = v.scale;
    sh[2 * w + 1] = v.ssq;
  }
  __syncthreads();
  Ssq r = ssq_init();
  if (threadIdx.x == 0)
    for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i) r = ssq_merge(r, Ssq{sh[2 * i], sh[2 * i + 1]});
  __syncthreads();
  return r;
}

}  // namespace mpir
