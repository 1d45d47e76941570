// rp_nm.hip -- the dense FP64 normal-mode transform on MFMA (plain path).
//
// Y = T X for a set of n-vectors (the reference's dsymv calls, verletmodule.f90:259,
// 279).  One wavefront computes one 16x16 output tile with V_MFMA_F64_16X16X4_F64:
//   A (16x4): lane l holds A[l%16][l/16]   = T[i0 + l%16][k0 + l/16]
//   B (4x16): lane l holds B[l/16][l%16]   = X[col c0 + l%16][bead k0 + l/16]
//   D (16x16): lane l holds D[l/16 + 4*j][l%16], j = 0..3 (checked on gfx950: item j of
//   a lane group is row 4j+g, so accumulator item j is directly the B operand of K rows 4j..4j+3)
// T is symmetric, so A is read as T[k][i] rows: 16 consecutive doubles per K row.
#include "rp_nm.h"

namespace rp {

__global__ void __launch_bounds__(64) nm_gemm_kernel(const double* __restrict__ T, int n, int ncols,
                                                     const double* __restrict__ in,
                                                     double* __restrict__ out,
                                                     const int* __restrict__ shift,
                                                     const double* __restrict__ bv, int backward) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
  const int m = lane & 15, kq = lane >> 4;
  const int col = c0 + m;
  const int sh = (col < ncols && shift) ? shift[col] : -1;
  const int row_a = i0 + m;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < n; k0 += 4) {
    const int k = k0 + kq;
    double a = (row_a < n && k < n) ? T[(size_t)k * n + row_a] : 0.0;
    double b = 0.0;
    if (col < ncols && k < n) {
      b = in[(size_t)col * n + k];
      if (backward && sh >= 0) b = b + bv[(size_t)sh * n + k];
    }
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  if (col < ncols) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = i0 + kq + 4 * j;
      if (row < n) {
        double v = acc[j];
        if (!backward && sh >= 0) v = v - bv[(size_t)sh * n + row];
        out[(size_t)col * n + row] = v;
      }
    }
  }
}

int launch_nm_gemm(const double* T, int n, int ncols, const double* in, double* out,
                   const int* shift, const double* bv, int backward, hipStream_t s) {
  RP_REQUIRE(n > 0 && ncols > 0, "nm transform of n=%d ncols=%d", n, ncols);
  dim3 grid((n + 15) / 16, (ncols + 15) / 16);
  nm_gemm_kernel<<<grid, 64, 0, s>>>(T, n, ncols, in, out, shift, bv, backward);
  RP_HIP(hipGetLastError());
  return RP_OK;
}

}  // namespace rp
