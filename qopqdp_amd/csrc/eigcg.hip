// Block operations of eigCG on parity-half vectors: rotations by a small
// dense matrix, y_j = sum_i R(i, j) x_i (lib/invert_eigcg_p.c:105-187), and
// Gram matrices X^H Y (its Rayleigh-Ritz and Lanczos matrices and
// Gram-Schmidt projections).  Rotation: one thread per complex element keeps
// 8 outputs in registers and streams the inputs, so the n1 inputs are read
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

// Gram matrix of two blocks of parity-half vectors stored with stride ld:
// G(i, j) = sum_k conj(X_i[k]) Y_j[k], i < n1, j < n2 (n1 n2 <= 16 * 256,
// (n1 + n2) tiles of GT + 1 complex in LDS).  Each block reduces one chunk
// of k through LDS tiles (rows padded by one element against bank
// conflicts) into per-block partial sums; gram_reduce_kernel adds them in
// block order.
constexpr int GT = 32;
constexpr int GP = 16;

__global__ __launch_bounds__(256) void gram_kernel(const double2 *X, const double2 *Y, long ld, int n1,
                                                   int n2, long K, long chunk, double2 *partial) {
  extern __shared__ double2 sh[];
  double2 *Xs = sh, *Ys = sh + n1 * (GT + 1);
  const long kb = blockIdx.x * chunk;
  const long ke = kb + chunk < K ? kb + chunk : K;
  const int np = n1 * n2, tid = threadIdx.x;
  double2 acc[GP];
#pragma unroll
  for (int q = 0; q < GP; q++) acc[q] = make_double2(0, 0);
  for (long k0 = kb; k0 < ke; k0 += GT) {
    for (int idx = tid; idx < (n1 + n2) * GT; idx += 256) {
      const int v = idx / GT, kk = idx - v * GT;
      const long k = k0 + kk;
      double2 val = make_double2(0, 0);
      if (k < ke) val = v < n1 ? X[(long)v * ld + k] : Y[(long)(v - n1) * ld + k];
      sh[v * (GT + 1) + kk] = val;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < GP; q++) {
      const int p = tid + q * 256;
      if (p < np) {
        const int i = p / n2, j = p - i * n2;
        const double2 *xr = Xs + i * (GT + 1), *yr = Ys + j * (GT + 1);
        double re = acc[q].x, im = acc[q].y;
        for (int kk = 0; kk < GT; kk++) {
          const double2 a = xr[kk], b = yr[kk];
          re = fma(a.x, b.x, fma(a.y, b.y, re));
          im = fma(a.x, b.y, fma(-a.y, b.x, im));
        }
        acc[q] = make_double2(re, im);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < GP; q++) {
    const int p = tid + q * 256;
    if (p < np) partial[(long)blockIdx.x * np + p] = acc[q];
  }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(const double2 *partial, int nblk, int np,
                                                          double2 *G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double re = 0, im = 0;
  for (int b = 0; b < nblk; b++) {
    re += partial[(long)b * np + p].x;
    im += partial[(long)b * np + p].y;
  }
  G[p] = make_double2(re, im);
}

}  // namespace

int gram_max_pairs() { return GP * 256; }
int gram_max_rows() { return (64 * 1024) / ((GT + 1) * (int)sizeof(double2)); }

void launch_gram(const double2 *X, const double2 *Y, long ld, int n1, int n2, long K, double2 *partial,
                 int nblk, double2 *G, hipStream_t st) {
  if (n1 <= 0 || n2 <= 0) return;
  const long chunk = ((K + nblk - 1) / nblk + GT - 1) / GT * GT;
  const size_t lds = (size_t)(n1 + n2) * (GT + 1) * sizeof(double2);
  hipLaunchKernelGGL(gram_kernel, dim3(nblk), dim3(256), lds, st, X, Y, ld, n1, n2, K, chunk, partial);
  const int np = n1 * n2;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((np + 255) / 256), dim3(256), 0, st, partial, nblk, np, G);
  QOP_CHECK_HIP(hipGetLastError());
}

void launch_rotate(const double2 *const *x, double2 *const *y, const double2 *R, int n1, int n2,
                   int ld, long len, hipStream_t st) {
  if (n1 <= 0 || n2 <= 0 || len <= 0) return;
  hipLaunchKernelGGL(rotate_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, x, y, R, n1,
                     n2, ld, len);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
