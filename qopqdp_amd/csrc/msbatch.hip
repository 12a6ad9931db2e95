// Device kernels of the batched multi-shift CG (msbatch.h).  The scalar
// recurrences follow lib/invert_cg_p.c:203-246 and :349-376 per source.
#include "msbatch.h"

namespace qop {

template <int P> struct MT;
template <> struct MT<P32> { using R = float; using S = float2; };
template <> struct MT<P16> { using R = float; using S = __half2; };

__device__ __forceinline__ double msum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

constexpr int MS_THREADS = 1024;
__device__ double ms_block_sum(const double *p, int m, double *sh) {
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int i = threadIdx.x;
  for (; i + 3 * MS_THREADS < m; i += 4 * MS_THREADS) {
    s0 += p[i];
    s1 += p[i + MS_THREADS];
    s2 += p[i + 2 * MS_THREADS];
    s3 += p[i + 3 * MS_THREADS];
  }
  for (; i < m; i += MS_THREADS) s0 += p[i];
  double s = msum((s0 + s1) + (s2 + s3));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = 0;
  if (threadIdx.x < 64) r = msum(threadIdx.x < MS_THREADS / 64 ? sh[threadIdx.x] : 0.0);
  if (threadIdx.x == 0) sh[16] = r;
  __syncthreads();
  r = sh[16];
  __syncthreads();
  return r;
}

// pAp, beta_0 and the zeta/beta recurrences of the live shifts
__global__ void __launch_bounds__(1024)
ms_scalarA(MsSlot *S, int *active, const double *partA, const double *partB, int nch) {
  __shared__ double sh[17];
  const int b = blockIdx.x;
  if (!active[b]) return;
  double Ain2 = ms_block_sum(partA + (long)b * nch, nch, sh);
  double p2 = ms_block_sum(partB + (long)b * nch, nch, sh);
  if (threadIdx.x != 0) return;
  MsSlot &s = S[b];
  if (s.done_pending) {  // the previous iteration was the last one
    s.done = 1;
    active[b] = 0;
    return;
  }
  for (int i = 0; i < s.ns; i++) s.live[i] = s.live_next[i];
  const int imin = s.imin;
  s.r2_old = s.r2;
  const double pAp = Ain2 + s.m2 * p2;
  s.pAp = pAp;
  s.iter++;
  if (!(pAp > 0)) {  // loss of precision (invert_cg_p.c:201)
    s.beta0 = 0;
    for (int i = 0; i < s.ns; i++) { s.b[i] = 0; s.live[i] = 0; s.live_next[i] = 0; }
    s.done_pending = 1;
    return;
  }
  const double b0 = s.r2 / pAp;
  s.b[imin] = b0;
  s.beta0 = b0;
  s.zn[imin] = 1;
  for (int i = 0; i < s.ns; i++) {
    if (i == imin) continue;
    if (!s.live[i]) { s.b[i] = 0; continue; }
    double zn = s.z[i] * s.zo[i] * s.bo[imin];
    double c1 = b0 * s.a[imin] * (s.zo[i] - s.z[i]);
    c1 += s.zo[i] * s.bo[imin] * (1 + s.shift[i] * b0);
    if (c1 != 0.0) zn /= c1; else zn = 0;
    if (s.z[i] != 0.0) s.b[i] = b0 * zn / s.z[i];
    else { zn = 0; s.b[i] = 0; }
    s.zn[i] = zn;
  }
}

// new |r|^2, alpha, per-shift stopping, shift-over of zeta/beta
__global__ void __launch_bounds__(1024)
ms_scalarB(MsSlot *S, int *active, const double *partC, int nch, int *ru) {
  __shared__ double sh[17];
  const int b = blockIdx.x;
  if (!active[b]) return;
  double r2 = ms_block_sum(partC + (long)b * nch, nch, sh);
  if (threadIdx.x != 0) return;
  MsSlot &s = S[b];
  if (s.done_pending) return;
  const int imin = s.imin;
  s.r2 = r2;
  const double a0 = r2 / s.r2_old;
  s.a[imin] = a0;
  int any = 0;
  for (int i = 0; i < s.ns; i++) {
    if (i != imin) {
      double c2 = s.z[i] * s.b[imin];
      s.a[i] = c2 != 0.0 ? a0 * s.zn[i] * s.b[i] / c2 : 0;
    }
    int ln = s.live[i] && !(s.zn[i] * s.zn[i] * r2 <= s.stop2[i]);
    s.live_next[i] = ln;
    any |= ln;
  }
  for (int i = 0; i < s.ns; i++) { s.bo[i] = s.b[i]; s.zo[i] = s.z[i]; s.z[i] = s.zn[i]; }
  if (!any || s.iter >= s.max_iter) s.done_pending = 1;
  // reliable update once the residual has dropped by delta since the last one
  if (!s.done_pending && r2 < s.delta2 * s.r2_max) ru[b] = 1;
}

// x64[k] += xs[k], xs[k] = 0 for every shift of a flagged slot
__global__ void __launch_bounds__(256)
ms_ru_fold(const int *ru, int Vh, VecView x64, VecView xs, long stride, int cap) {
  const int k = blockIdx.y;
  if (!ru[k % cap]) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= Vh) return;
  double2 *x = (double2 *)x64.data + (size_t)k * stride;
  float2 *y = (float2 *)xs.data + (size_t)k * stride;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float2 v = y[c * Vh + h];
    double2 t = x[c * Vh + h];
    t.x += (double)v.x;
    t.y += (double)v.y;
    x[c * Vh + h] = t;
    y[c * Vh + h] = make_float2(0.f, 0.f);
  }
}

// r = b - ax (FP64) stored as the sloppy residual, |r|^2 partials
template <int P>
__global__ void __launch_bounds__(256)
ms_ru_resid(const int *ru, int Vh, VecView b64, VecView ax64, VecView rs, long stride,
            double *part) {
  using R = typename MT<P>::R;
  const int b = blockIdx.y;
  if (!ru[b]) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  if (h < Vh) {
    const double2 *bb = (const double2 *)b64.data + (size_t)b * stride;
    const double2 *aa = (const double2 *)ax64.data + (size_t)b * stride;
    R r[6];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double2 u = bb[c * Vh + h], w = aa[c * Vh + h];
      const double xr = u.x - w.x, xi = u.y - w.y;
      acc += xr * xr + xi * xi;
      r[2 * c] = (R)xr;
      r[2 * c + 1] = (R)xi;
    }
    const size_t cb = sizeof(typename MT<P>::S);
    char *d = (char *)rs.data + (size_t)b * stride * cb;
    if constexpr (P == P16) {
      float *nm = rs.norm + (long)b * Vh;
      float m = 0.f;
#pragma unroll
      for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(r[i]));
      const float inv = m > 0.f ? 1.f / m : 0.f;
#pragma unroll
      for (int c = 0; c < 3; c++)
        ((__half2 *)d)[c * Vh + h] = __floats2half2_rn(r[2 * c] * inv, r[2 * c + 1] * inv);
      nm[h] = m;
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) ((float2 *)d)[c * Vh + h] = make_float2(r[2 * c], r[2 * c + 1]);
    }
  }
  const int nch = (Vh + 63) / 64;
  acc = msum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) part[(long)b * nch + (h >> 6)] = acc;
}

// the true |r|^2 replaces the iterated one
__global__ void __launch_bounds__(1024)
ms_ru_scalar(MsSlot *S, int *ru, const double *part, int nch) {
  __shared__ double sh[17];
  const int b = blockIdx.x;
  if (!ru[b]) return;
  const double r2 = ms_block_sum(part + (long)b * nch, nch, sh);
  if (threadIdx.x != 0) return;
  MsSlot &s = S[b];
  s.r2 = r2;
  s.r2_max = r2;
  s.nupdate++;
  ru[b] = 0;
}

// x_i += b_i p_i (old p_i), p_i = zn_i r + a_i p_i (live shifts), p = r + a_0 p,
// partial |p|^2.  The base solution was updated in the hop epilogue.
template <int P>
__global__ void __launch_bounds__(256)
ms_update(const MsSlot *S, const int *active, int Vh, VecView rs, VecView ps, VecView pm,
          VecView xm, long stride, long sstride, int cap, double *partB) {
  using R = typename MT<P>::R;
  using Sx = typename MT<P>::S;
  const int b = blockIdx.y;
  if (!active[b]) return;
  const MsSlot &s = S[b];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  if (h < Vh) {
    const size_t cb = sizeof(Sx);
    const size_t off = (size_t)b * stride;
    const long noff = (long)b * Vh;
    R r[6], p[6];
    auto ld = [&](const VecView &v, size_t o, const float *nm, R *out) {
      if constexpr (P == P16) {
        const float n = nm[h];
        for (int c = 0; c < 3; c++) {
          float2 t = __half22float2(((const __half2 *)((const char *)v.data + o * cb))[c * Vh + h]);
          out[2 * c] = t.x * n; out[2 * c + 1] = t.y * n;
        }
      } else {
        for (int c = 0; c < 3; c++) {
          float2 t = ((const float2 *)((const char *)v.data + o * cb))[c * Vh + h];
          out[2 * c] = t.x; out[2 * c + 1] = t.y;
        }
      }
    };
    auto st_ = [&](const VecView &v, size_t o, float *nm, R *in) {
      if constexpr (P == P16) {
        float m = 0.f;
        for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(in[i]));
        float inv = m > 0.f ? 1.f / m : 0.f;
        for (int c = 0; c < 3; c++) {
          // products and packed halves formed once (empty asm), so the
          // values kept are the words stored
          float a0 = in[2 * c] * inv, a1 = in[2 * c + 1] * inv;
          asm volatile("" : "+v"(a0), "+v"(a1));
          unsigned qw = __builtin_bit_cast(unsigned, __floats2half2_rn(a0, a1));
          asm volatile("" : "+v"(qw));
          ((unsigned *)((char *)v.data + o * cb))[c * Vh + h] = qw;
          const float2 f = __half22float2(__builtin_bit_cast(__half2, qw));
          in[2 * c] = (R)f.x * m;
          in[2 * c + 1] = (R)f.y * m;
        }
        nm[h] = m;
      } else {
        for (int c = 0; c < 3; c++)
          ((float2 *)((char *)v.data + o * cb))[c * Vh + h] = make_float2(in[2 * c], in[2 * c + 1]);
      }
    };
    ld(rs, off, rs.norm ? rs.norm + noff : nullptr, r);
    for (int i = 0; i < s.ns; i++) {
      if (i == s.imin || !s.live[i]) continue;
      const size_t o = (size_t)(i * cap + b) * stride;
      const long no = (long)(i * cap + b) * Vh;
      R q[6];
      float *qn = pm.norm ? pm.norm + no : nullptr;
      ld(pm, o, qn, q);
      // x_i (FP32) += b_i p_i
      float2 *x = (float2 *)xm.data + (size_t)(i * cap + b) * stride;
      const float bi = (float)s.b[i];
      for (int c = 0; c < 3; c++) {
        float2 t = x[c * Vh + h];
        t.x = fmaf(bi, (float)q[2 * c], t.x);
        t.y = fmaf(bi, (float)q[2 * c + 1], t.y);
        x[c * Vh + h] = t;
      }
      if (s.live_next[i] && !s.done_pending) {
        const R zi = (R)s.zn[i], ai = (R)s.a[i];
        for (int e = 0; e < 6; e++) q[e] = fma(ai, q[e], zi * r[e]);
        st_(pm, o, qn, q);
      }
    }
    if (!s.done_pending) {
      float *pn = ps.norm ? ps.norm + noff : nullptr;
      ld(ps, off, pn, p);
      const R a0 = (R)s.a[s.imin];
      for (int e = 0; e < 6; e++) p[e] = fma(a0, p[e], r[e]);
      st_(ps, off, pn, p);
      for (int e = 0; e < 6; e++) acc += (double)p[e] * (double)p[e];
    }
  }
  const int nch = (Vh + 63) / 64;
  acc = msum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) partB[(long)b * nch + (h >> 6)] = acc;
  (void)sstride;
}

void ms_launch_scalarA(MsSlot *s, const int *active, const double *partA, const double *partB,
                       int nch, int n, hipStream_t st) {
  hipLaunchKernelGGL(ms_scalarA, dim3(n), dim3(MS_THREADS), 0, st, s, const_cast<int *>(active), partA, partB, nch);
  QOP_CHECK_HIP(hipGetLastError());
}
void ms_launch_scalarB(MsSlot *s, int *active, const double *partC, int nch, int n, int *ru,
                       hipStream_t st) {
  hipLaunchKernelGGL(ms_scalarB, dim3(n), dim3(MS_THREADS), 0, st, s, active, partC, nch, ru);
  QOP_CHECK_HIP(hipGetLastError());
}
void ms_launch_ru_fold(const int *ru, int Vh, VecView x64, VecView xs, long stride, int cap, int ns,
                       hipStream_t st) {
  hipLaunchKernelGGL(ms_ru_fold, dim3((Vh + 255) / 256, cap * ns), dim3(256), 0, st, ru, Vh, x64, xs,
                     stride, cap);
  QOP_CHECK_HIP(hipGetLastError());
}
void ms_launch_ru_resid(const int *ru, int prec, int Vh, VecView b64, VecView ax64, VecView rs,
                        long stride, int cap, double *part, hipStream_t st) {
  dim3 g((Vh + 255) / 256, cap), bl(256);
  if (prec == P16) hipLaunchKernelGGL(ms_ru_resid<P16>, g, bl, 0, st, ru, Vh, b64, ax64, rs, stride, part);
  else hipLaunchKernelGGL(ms_ru_resid<P32>, g, bl, 0, st, ru, Vh, b64, ax64, rs, stride, part);
  QOP_CHECK_HIP(hipGetLastError());
}
void ms_launch_ru_scalar(MsSlot *s, int *ru, const double *part, int nch, int n, hipStream_t st) {
  hipLaunchKernelGGL(ms_ru_scalar, dim3(n), dim3(MS_THREADS), 0, st, s, ru, part, nch);
  QOP_CHECK_HIP(hipGetLastError());
}
void ms_launch_update(const MsSlot *s, const int *active, int prec, int Vh, VecView rs, VecView ps,
                      VecView pm, VecView xm, long stride, long sstride, int cap, int ns, double *partB,
                      hipStream_t st) {
  dim3 g((Vh + 255) / 256, cap), bl(256);
  if (prec == P16) hipLaunchKernelGGL(ms_update<P16>, g, bl, 0, st, s, active, Vh, rs, ps, pm, xm, stride, sstride, cap, partB);
  else hipLaunchKernelGGL(ms_update<P32>, g, bl, 0, st, s, active, Vh, rs, ps, pm, xm, stride, sstride, cap, partB);
  QOP_CHECK_HIP(hipGetLastError());
  (void)ns;
}

}  // namespace qop
