// Rotation of a set of parity-half vectors by a small dense matrix,
// y_j = sum_i R(i, j) x_i (the vector rotations of eigCG,
// lib/invert_eigcg_p.c:105-187).  One thread per complex element keeps 8
// outputs in registers and streams the inputs, so the n1 inputs are read
// ceil(n2 / 8) times; R is uniform across the grid.
#include "eigcg.h"

namespace qop {
namespace {

constexpr int ROT_J = 8;

__global__ __launch_bounds__(256) void rotate_kernel(const double2 *const *x, double2 *const *y,
                                                     const double2 *R, int n1, int n2, int ld,
                                                     long len) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= len) return;
  for (int j0 = 0; j0 < n2; j0 += ROT_J) {
    double2 acc[ROT_J];
#pragma unroll
    for (int jj = 0; jj < ROT_J; jj++) acc[jj] = make_double2(0, 0);
    for (int i = 0; i < n1; i++) {
      const double2 xi = x[i][k];
      const double2 *r = R + (long)i * ld + j0;
#pragma unroll
      for (int jj = 0; jj < ROT_J; jj++) {
        if (j0 + jj < n2) {
          const double2 c = r[jj];
          acc[jj].x = fma(c.x, xi.x, fma(-c.y, xi.y, acc[jj].x));
          acc[jj].y = fma(c.x, xi.y, fma(c.y, xi.x, acc[jj].y));
        }
      }
    }
#pragma unroll
    for (int jj = 0; jj < ROT_J; jj++)
      if (j0 + jj < n2) y[j0 + jj][k] = acc[jj];
  }
}

}  // namespace

void launch_rotate(const double2 *const *x, double2 *const *y, const double2 *R, int n1, int n2,
                   int ld, long len, hipStream_t st) {
  if (n1 <= 0 || n2 <= 0 || len <= 0) return;
  hipLaunchKernelGGL(rotate_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, x, y, R, n1,
                     n2, ld, len);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
