// Device-side pieces of the multi-shift CG (lib/invert_cg_p.c:130-402):
// the per-iteration scalar recurrences (zeta, beta, alpha in FP64) run in a
// one-block kernel, and the vector updates of all shifts are fused into two
// passes.  The host only intervenes for the solution fold / true-residual
// check, which the reference does every 100 iterations or on convergence.
#include "solver.h"

namespace qop {

template <int P> struct QT;
template <> struct QT<P64> { using R = double; using S = double2; };
template <> struct QT<P32> { using R = float; using S = float2; };

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum m partials in the same fixed order as finalize_kernel (256 threads).
__device__ double block_sum256(const double *p, int m, double *sh) {
  double s = 0;
  for (int i = threadIdx.x; i < m; i += 256) s += p[i];
  s = wsum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// After linop: pkp, b[imin] and the zeta/beta recurrences (:203-238).
__global__ void __launch_bounds__(256)
cgms_scalarA(CgmsDev d, int nch) {
  __shared__ double sh[4];
  CgmsState *s = d.st;
  if (s->status != 0) return;
  double Ain2 = block_sum256(d.partA, nch, sh);
  double in2 = block_sum256(d.partB, nch, sh);
  if (threadIdx.x != 0) return;
  s->oldrsq = s->rsq;
  double pkp = Ain2 + s->m2 * in2;  // ddagd_norm2 return value
  s->pkp = pkp;
  s->iteration++;
  if (pkp <= 0) { s->status = 2; s->break_reason = 1; return; }
  const int imin = s->imin, ns = s->ns;
  s->b[imin] = s->rsq / pkp;
  s->zn[imin] = 1;
  for (int i = 0; i < ns; i++) {
    if (i == imin) continue;
    double zn = s->z[i] * s->zo[i] * s->bo[imin];
    double c1 = s->b[imin] * s->a[imin] * (s->zo[i] - s->z[i]);
    c1 += s->zo[i] * s->bo[imin] * (1 + s->shifts[i] * s->b[imin]);
    if (c1 != 0.0) zn /= c1; else zn = 0;
    if (s->z[i] != 0.0) s->b[i] = s->b[imin] * zn / s->z[i];
    else { zn = 0; s->b[i] = 0; }
    s->zn[i] = zn;
  }
}

// r -= b[imin] Mp ; x_i += b_i p_i ; partial |r|^2   (:209-252)
template <int P>
__global__ void __launch_bounds__(256)
cgms_update(CgmsDev d, int Vh, VecView r, VecView Mp, int ns) {
  using R = typename QT<P>::R;
  using S = typename QT<P>::S;
  const CgmsState *s = d.st;
  if (s->status != 0) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  if (h < Vh) {
    const R b0 = (R)s->b[s->imin];
    S *rr = (S *)r.data;
    const S *mp = (const S *)Mp.data;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      S a = rr[c * Vh + h], m = mp[c * Vh + h];
      a.x = fma(-b0, m.x, a.x);
      a.y = fma(-b0, m.y, a.y);
      rr[c * Vh + h] = a;
      acc += (double)a.x * a.x + (double)a.y * a.y;
    }
    for (int i = 0; i < ns; i++) {
      const R bi = (R)s->b[i];
      S *x = (S *)d.xs[i].data;
      const S *pm = (const S *)d.pms[i].data;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        S xv = x[c * Vh + h], pv = pm[c * Vh + h];
        xv.x = fma(bi, pv.x, xv.x);
        xv.y = fma(bi, pv.y, xv.y);
        x[c * Vh + h] = xv;
      }
    }
  }
  const int nch = (Vh + 63) / 64;
  acc = wsum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) d.partC[h >> 6] = acc;
}

// Relative-residual reduction (only when relmin > 0): r vs out[irel].
template <int P>
__global__ void __launch_bounds__(256)
cgms_relnorm(CgmsDev d, int Vh, VecView r, VecView x) {
  using S = typename QT<P>::S;
  const CgmsState *s = d.st;
  if (s->status != 0) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  if (h < Vh) {
    double r2 = 0, x2 = 0;
    for (int c = 0; c < 3; c++) {
      S a = ((const S *)r.data)[c * Vh + h], b = ((const S *)x.data)[c * Vh + h];
      r2 += (double)a.x * a.x + (double)a.y * a.y;
      x2 += (double)b.x * b.x + (double)b.y * b.y;
    }
    acc = x2 == 0.0 ? 1.0 : r2 / x2;
  }
  const int nch = (Vh + 63) / 64;
  acc = wsum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) d.partD[h >> 6] = acc;
}

// After the update: rsq, stopping tests, then alpha and the shift-over of
// zeta/beta (:254-275, 349-376).  Sets status 1 when the host must fold the
// solution and check the true residual.
__global__ void __launch_bounds__(256)
cgms_scalarB(CgmsDev d, int nch, int with_rel, int Vh) {
  __shared__ double sh[4];
  CgmsState *s = d.st;
  if (s->status != 0) return;
  double rsq = block_sum256(d.partC, nch, sh);
  double rel = with_rel ? block_sum256(d.partD, nch, sh) : 0.0;
  if (threadIdx.x != 0) return;
  s->rsq = rsq;
  const int it = s->iteration;
  double rsqc = s->zn[s->irsq] * s->zn[s->irsq] * rsq;
  if (with_rel) s->relnorm2 = s->zn[s->irel] * s->zn[s->irel] * (rel / (double)Vh);
  if (it >= s->restart || it >= s->max_iter) { s->status = 2; s->break_reason = 2; return; }
  if (rsqc * s->trueRsqFac < s->rsqstop) s->nextRsqCheck = it;
  if (with_rel && s->relnorm2 * s->trueRelFac < s->relstop) s->nextRelCheck = it;
  if (s->nextRsqCheck <= it || (s->relstop > 0 && s->nextRelCheck <= it)) {
    s->status = 1;
    return;
  }
  // alpha and shift-over
  const int imin = s->imin, ns = s->ns;
  s->a[imin] = rsq / s->oldrsq;
  for (int i = 0; i < ns; i++) {
    if (i == imin) continue;
    double c2 = s->z[i] * s->b[imin];
    s->a[i] = c2 != 0.0 ? s->a[imin] * s->zn[i] * s->b[i] / c2 : 0;
  }
  for (int i = 0; i < ns; i++) { s->bo[i] = s->b[i]; s->zo[i] = s->z[i]; s->z[i] = s->zn[i]; }
}

// p = r + a[imin] p ; p_i = zn_i r + a_i p_i   (:349-376); pms[imin] is p
template <int P>
__global__ void __launch_bounds__(256)
cgms_pupdate(CgmsDev d, int Vh, VecView r, int ns) {
  using R = typename QT<P>::R;
  using S = typename QT<P>::S;
  const CgmsState *s = d.st;
  if (s->status != 0) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= Vh) return;
  S rv[3];
  const S *rr = (const S *)r.data;
#pragma unroll
  for (int c = 0; c < 3; c++) rv[c] = rr[c * Vh + h];
  for (int i = 0; i < ns; i++) {
    const bool base = i == s->imin;
    const R ai = (R)s->a[i], zi = (R)s->zn[i];
    S *pm = (S *)d.pms[i].data;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      S pv = pm[c * Vh + h];
      if (base) {
        pv.x = fma(ai, pv.x, rv[c].x);
        pv.y = fma(ai, pv.y, rv[c].y);
      } else {
        R tx = zi * rv[c].x, ty = zi * rv[c].y;
        pv.x = fma(ai, pv.x, tx);
        pv.y = fma(ai, pv.y, ty);
      }
      pm[c * Vh + h] = pv;
    }
  }
}

void cgms_launch_scalarA(CgmsDev d, int nch, hipStream_t st) {
  hipLaunchKernelGGL(cgms_scalarA, dim3(1), dim3(256), 0, st, d, nch);
  QOP_CHECK_HIP(hipGetLastError());
}
void cgms_launch_update(CgmsDev d, int prec, int Vh, VecView r, VecView Mp, int ns,
                        hipStream_t st) {
  dim3 g((Vh + 255) / 256), b(256);
  if (prec == P64) hipLaunchKernelGGL(cgms_update<P64>, g, b, 0, st, d, Vh, r, Mp, ns);
  else hipLaunchKernelGGL(cgms_update<P32>, g, b, 0, st, d, Vh, r, Mp, ns);
  QOP_CHECK_HIP(hipGetLastError());
}
void cgms_launch_relnorm(CgmsDev d, int prec, int Vh, VecView r, VecView x,
                         hipStream_t st) {
  dim3 g((Vh + 255) / 256), b(256);
  if (prec == P64) hipLaunchKernelGGL(cgms_relnorm<P64>, g, b, 0, st, d, Vh, r, x);
  else hipLaunchKernelGGL(cgms_relnorm<P32>, g, b, 0, st, d, Vh, r, x);
  QOP_CHECK_HIP(hipGetLastError());
}
void cgms_launch_scalarB(CgmsDev d, int nch, int with_rel, int Vh, hipStream_t st) {
  hipLaunchKernelGGL(cgms_scalarB, dim3(1), dim3(256), 0, st, d, nch, with_rel, Vh);
  QOP_CHECK_HIP(hipGetLastError());
}
void cgms_launch_pupdate(CgmsDev d, int prec, int Vh, VecView p, VecView r, int ns,
                         hipStream_t st) {
  (void)p;
  dim3 g((Vh + 255) / 256), b(256);
  if (prec == P64) hipLaunchKernelGGL(cgms_pupdate<P64>, g, b, 0, st, d, Vh, r, ns);
  else hipLaunchKernelGGL(cgms_pupdate<P32>, g, b, 0, st, d, Vh, r, ns);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
