// HISQ link preparation on the device: U(3) projection of the fat7-smeared
// links, W = V (V^dagger V)^-1/2, one thread per link, FP64.
//
// The analytic path follows QOP_u3_un_analytic
// (lib/hisq_reunit_utilities_p.c:97-297): the eigenvalues g0..g2 of
// Q = V^dagger V from the traces of Q, Q^2, Q^3 (cubic in trigonometric
// form), then Q^-1/2 = f0 + f1 Q + f2 Q^2 with the symmetric-function
// coefficients, and the same tests that make the reference fall back to its
// SVD (|det V|^2 against g0 g1 g2, |det V|^2 below an absolute bound, or the
// fallback forced).  The fallback here is the scaled Newton iteration for the
// unitary polar factor, X <- (z X + (z X)^-dagger) / 2 with z = |det X|^-1/3,
// which converges to the same matrix U_left V_right^dagger the reference's
// SVD gives; it runs at most 40 iterations.
#include "common.h"
#include "hisq.h"

namespace qop {
namespace {

struct C2 {
  double r, i;
};
__device__ __forceinline__ C2 cmul(C2 a, C2 b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ C2 cadd(C2 a, C2 b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ C2 csub(C2 a, C2 b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ C2 conjg(C2 a) { return {a.r, -a.i}; }

struct M3 {
  C2 e[3][3];
};

__device__ __forceinline__ void load_m(const double *p, M3 &m) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[i][j] = {p[6 * i + 2 * j], p[6 * i + 2 * j + 1]};
}
__device__ __forceinline__ void store_m(double *p, const M3 &m) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      p[6 * i + 2 * j] = m.e[i][j].r;
      p[6 * i + 2 * j + 1] = m.e[i][j].i;
    }
}
// c = a b
__device__ __forceinline__ void mul(M3 &c, const M3 &a, const M3 &b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; k++) s = cadd(s, cmul(a.e[i][k], b.e[k][j]));
      c.e[i][j] = s;
    }
}
// c = a^dagger b
__device__ __forceinline__ void mul_an(M3 &c, const M3 &a, const M3 &b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; k++) s = cadd(s, cmul(conjg(a.e[k][i]), b.e[k][j]));
      c.e[i][j] = s;
    }
}
__device__ __forceinline__ C2 det3(const M3 &a) {
  C2 d = {0, 0};
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    C2 minor = csub(cmul(a.e[1][j1], a.e[2][j2]), cmul(a.e[1][j2], a.e[2][j1]));
    d = cadd(d, cmul(a.e[0][j], minor));
  }
  return d;
}

// Unitary polar factor by the determinant-scaled Newton iteration.
__device__ bool polar_newton(const M3 &v, M3 &w) {
  M3 x = v;
#pragma unroll 1
  for (int it = 0; it < 40; it++) {
    C2 d = det3(x);
    const double ad = sqrt(d.r * d.r + d.i * d.i);
    if (!(ad > 0)) return false;  // singular: no polar factor from this iteration
    const double z = pow(ad, -1.0 / 3.0);
    // x^-dagger = adj(x)^dagger / conj(det x); adj(x)[i][j] = cofactor(x)[j][i]
    M3 xi;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        // cofactor C[i][j]
        xi.e[i][j] = csub(cmul(x.e[i1][j1], x.e[i2][j2]), cmul(x.e[i1][j2], x.e[i2][j1]));
      }
    // x^-dagger[i][j] = conj(x^-1[j][i]) = conj(C[i][j] / det) = conj(C[i][j]) / conj(det)
    const double inv = 1.0 / (ad * ad);
    const C2 dinv = {d.r * inv, d.i * inv};  // 1 / conj(det) = det / |det|^2
    double diff = 0, norm = 0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const C2 t = cmul(conjg(xi.e[i][j]), dinv);
        const C2 n = {0.5 * (z * x.e[i][j].r + t.r / z), 0.5 * (z * x.e[i][j].i + t.i / z)};
        diff += (n.r - x.e[i][j].r) * (n.r - x.e[i][j].r) + (n.i - x.e[i][j].i) * (n.i - x.e[i][j].i);
        norm += n.r * n.r + n.i * n.i;
        x.e[i][j] = n;
      }
    if (diff <= 1e-30 * norm) break;
  }
  w = x;
  return true;
}

__global__ void __launch_bounds__(256) u3_project_kernel(U3Args a) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long n = 4 * a.V;
  int flag = 0, bad = 0;
  if (t < n) {
    const int mu = (int)(t / a.V);
    const long s = t - mu * a.V;
    M3 v, w;
    load_m(a.in[mu] + 18 * s, v);
    const double eps = 1.0e-14;
    double det_check = 0;
    if (a.allow_svd) {
      C2 d = det3(v);
      det_check = d.r * d.r + d.i * d.i;
    }
    M3 q, q2, q3;
    mul_an(q, v, v);
    mul(q2, q, q);
    mul(q3, q2, q);
    const double c0 = q.e[0][0].r + q.e[1][1].r + q.e[2][2].r;
    const double c1 = (q2.e[0][0].r + q2.e[1][1].r + q2.e[2][2].r) / 2;
    const double c2 = (q3.e[0][0].r + q3.e[1][1].r + q3.e[2][2].r) / 3;
    double S = c1 / 3 - c0 * (c0 / 18);
    double g0, g1, g2;
    if (fabs(S) < eps) {
      g0 = g1 = g2 = c0 / 3;
    } else {
      const double R = c2 / 2 - c0 * (c1 / 3) + c0 * c0 * (c0 / 27);
      S = sqrt(S);
      const double RoS = R / (S * S * S);
      double theta;
      if (!(fabs(RoS) < 1.0)) theta = R > 0 ? 0.0 : M_PI;
      else theta = acos(RoS);
      const double th3 = theta / 3, p23 = 2 * M_PI / 3;
      g0 = c0 / 3 + 2 * S * cos(th3);
      g1 = c0 / 3 + 2 * S * cos(th3 + p23);
      g2 = c0 / 3 + 2 * S * cos(th3 + 2 * p23);
    }
    if (a.allow_svd) {
      if (!a.svd_only) {
        if (det_check != 0 && fabs(det_check - g0 * g1 * g2) / fabs(det_check) > a.rel_error) flag = 1;
        if (det_check < a.abs_error) flag = 1;
      } else {
        flag = 1;
      }
    }
    if (flag) {
      if (!polar_newton(v, w)) {
        bad = 1;
        w = v;
      }
    } else {
      const double g0s = sqrt(g0), g1s = sqrt(g1), g2s = sqrt(g2);
      double us = g1s + g2s, ws = g1s * g2s;
      const double vs = g0s * us + ws;
      us += g0s;
      ws *= g0s;
      const double den = ws * (us * vs - ws);
      const double f0 = (us * vs * vs - ws * (us * us + vs)) / den;
      const double f1 = (2 * us * vs - ws - us * us * us) / den;
      const double f2 = us / den;
      M3 sm;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
          sm.e[i][j] = {f1 * q.e[i][j].r + f2 * q2.e[i][j].r, f1 * q.e[i][j].i + f2 * q2.e[i][j].i};
#pragma unroll
      for (int i = 0; i < 3; i++) sm.e[i][i].r += f0;
      mul(w, v, sm);
    }
    store_m(a.out[mu] + 18 * s, w);
  }
  // per-wave counts of fallbacks and failures, written by lane 0 of each wave
  int cf = flag, cb = bad;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    cf += __shfl_xor(cf, off, 64);
    cb += __shfl_xor(cb, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    const long wave = t >> 6;
    a.counts[2 * wave] = cf;
    a.counts[2 * wave + 1] = cb;
  }
}

__global__ void axpy_raw_kernel(double *o, const double *a, double s, const double *b, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) o[i] = (a ? a[i] : 0.0) + s * b[i];
}

}  // namespace

long u3_nwaves(const Lattice *lat) { return (4 * lat->V + 255) / 256 * 4; }

void launch_u3_project(const U3Args &a, hipStream_t st) {
  const long n = 4 * a.V;
  const int nb = (int)((n + 255) / 256);
  u3_project_kernel<<<nb, 256, 0, st>>>(a);
  QOP_CHECK_HIP(hipGetLastError());
}

void launch_axpy_raw(double *o, const double *a, double s, const double *b, long n, hipStream_t st) {
  const int nb = (int)((n + 255) / 256);
  axpy_raw_kernel<<<nb, 256, 0, st>>>(o, a, s, b, n);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
