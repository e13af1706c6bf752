// FP64 sparse matrix-vector products for the refinement residual and the
// Arnoldi step.
//
// The reference multiplies with the upper CSC of the symmetric matrix
// (MatrixMultiply, include/MPIR/ops.hpp:25-45; ResidualCSCSymmetric,
// include/qdldl-luir/residual.hpp:14-44), scattering each off-diagonal entry
// to both triangles. Every output y_i then receives its terms in ascending
// column order starting from zero: first column i's entries (rows <= i,
// ascending), then the entries (i, j) of later columns j. A full CSR with
// ascending columns summed left to right from zero therefore produces the
// same rounding, so the kernels below keep that order and forbid fused
// multiply-add contraction: y is bit-identical to the reference's loop.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common.hpp"
#include "kernels.hpp"

namespace mpir {

namespace {

constexpr int kRowsPerBlock = 256;

// One thread per row. Column indices and values of the block's rows are
// staged through LDS in coalesced chunks; each thread then walks its row in
// order.
__global__ void __launch_bounds__(256) k_spmv(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cj,
                                              const double* __restrict__ av, const double* __restrict__ x,
                                              const double* __restrict__ b, double* __restrict__ y, int mode,
                                              const int* __restrict__ gate) {
#pragma clang fp contract(off)
  if (gate != nullptr && *gate == 0) return;
  constexpr int kChunk = 2048;
  __shared__ int32_t sj[kChunk];
  __shared__ double sv[kChunk];
  const int row0 = blockIdx.x * kRowsPerBlock;
  const int row1 = min(n, row0 + kRowsPerBlock);
  const int r = row0 + threadIdx.x;
  const int q0 = rp[row0], q1 = rp[row1];
  const int my0 = r < row1 ? rp[r] : q1;
  const int my1 = r < row1 ? rp[r + 1] : q1;
  double acc = 0.0;
  for (int c0 = q0; c0 < q1; c0 += kChunk) {
    const int c1 = min(q1, c0 + kChunk);
    __syncthreads();
    for (int q = c0 + threadIdx.x; q < c1; q += blockDim.x) {
      sj[q - c0] = cj[q];
      sv[q - c0] = av[q];
    }
    __syncthreads();
    const int lo = max(my0, c0), hi = min(my1, c1);
    for (int q = lo; q < hi; ++q) acc = acc + sv[q - c0] * x[sj[q - c0]];
  }
  if (r < row1) y[r] = mode ? b[r] - acc : acc;
}

}  // namespace

void spmv_f64(const CsrDev& A, const double* x, double* y, hipStream_t st) {
  if (A.n == 0) return;
  hipLaunchKernelGGL(k_spmv, dim3((A.n + kRowsPerBlock - 1) / kRowsPerBlock), dim3(kRowsPerBlock), 0, st, A.n, A.p,
                     A.j, A.x, x, nullptr, y, 0, A.gate);
  MPIR_HIP(hipGetLastError());
}

void residual_f64(const CsrDev& A, const double* x, const double* b, double* r, hipStream_t st) {
  if (A.n == 0) return;
  hipLaunchKernelGGL(k_spmv, dim3((A.n + kRowsPerBlock - 1) / kRowsPerBlock), dim3(kRowsPerBlock), 0, st, A.n, A.p,
                     A.j, A.x, x, b, r, 1, A.gate);
  MPIR_HIP(hipGetLastError());
}

}  // namespace mpir
