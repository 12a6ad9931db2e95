// Device kernels of the batched mixed-precision CG (see batch.h): per-slot
// scalar recurrences in FP64 and fused vector updates in the sloppy
// precision.  A slot's arithmetic never depends on other slots.
#include "batch.h"

namespace qop {

template <int P> struct BT;
template <> struct BT<P64> { using R = double; using S = double2; };
template <> struct BT<P32> { using R = float; using S = float2; };
template <> struct BT<P16> { using R = float; using S = __half2; };

__device__ __forceinline__ double bsum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Fixed-order sum of m partials by one 256-thread block.
__device__ double bs_block_sum(const double *p, int m, double *sh) {
  double s = 0;
  for (int i = threadIdx.x; i < m; i += 256) s += p[i];
  s = bsum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

template <int P>
__device__ __forceinline__ void ld3(const void *data, const float *norm, long h, long Vh,
                                    typename BT<P>::R v[6]) {
  if constexpr (P == P16) {
    const float n = norm[h];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float2 t = __half22float2(((const __half2 *)data)[c * Vh + h]);
      v[2 * c] = t.x * n; v[2 * c + 1] = t.y * n;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      typename BT<P>::S t = ((const typename BT<P>::S *)data)[c * Vh + h];
      v[2 * c] = t.x; v[2 * c + 1] = t.y;
    }
  }
}

template <int P>
__device__ __forceinline__ void st3(void *data, float *norm, long h, long Vh,
                                    const typename BT<P>::R v[6]) {
  if constexpr (P == P16) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(v[i]));
    float inv = m > 0.0f ? 1.0f / m : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++)
      ((__half2 *)data)[c * Vh + h] = __floats2half2_rn(v[2 * c] * inv, v[2 * c + 1] * inv);
    norm[h] = m;
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      typename BT<P>::S t;
      t.x = v[2 * c]; t.y = v[2 * c + 1];
      ((typename BT<P>::S *)data)[c * Vh + h] = t;
    }
  }
}

// pAp = |D p|^2 + m^2 |p|^2 ; alpha = r2 / pAp
__global__ void __launch_bounds__(256)
bs_scalar1(BsRhs *rhs, const int *active, const int *status, const double *partA,
           const double *partB, int nch) {
  __shared__ double sh[4];
  if (*status != 0) return;
  const int b = blockIdx.x;
  BsRhs &R = rhs[b];
  if (!active[b]) return;
  double a = bs_block_sum(partA + (long)b * nch, nch, sh);
  double p2 = bs_block_sum(partB + (long)b * nch, nch, sh);
  if (threadIdx.x != 0) return;
  double pAp = a + R.m2 * p2;
  R.pAp = pAp;
  R.alpha = pAp > 0 ? R.r2 / pAp : 0.0;
}

// rs -= alpha Aps ; xs += alpha ps ; partial |rs|^2
template <int P>
__global__ void __launch_bounds__(256)
bs_update(const BsRhs *rhs, const int *active, const int *status, int Vh, VecView rs,
          VecView Aps, VecView ps, VecView xs, long stride, double *partC) {
  using R = typename BT<P>::R;
  if (*status != 0) return;
  const int b = blockIdx.y;
  if (!active[b]) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  if (h < Vh) {
    const R al = (R)rhs[b].alpha;
    const size_t off = (size_t)b * stride;
    const long noff = (long)b * Vh;
    const size_t cb = sizeof(typename BT<P>::S);
    R r[6], a[6], p[6];
    ld3<P>((const char *)rs.data + off * cb, rs.norm ? rs.norm + noff : nullptr, h, Vh, r);
    ld3<P>((const char *)Aps.data + off * cb, Aps.norm ? Aps.norm + noff : nullptr, h, Vh, a);
    ld3<P>((const char *)ps.data + off * cb, ps.norm ? ps.norm + noff : nullptr, h, Vh, p);
#pragma unroll
    for (int i = 0; i < 6; i++) r[i] = fma(-al, a[i], r[i]);
    st3<P>((char *)rs.data + off * cb, rs.norm ? rs.norm + noff : nullptr, h, Vh, r);
    if constexpr (P == P16) {  // the norm uses the stored value
      ld3<P>((const char *)rs.data + off * cb, rs.norm + noff, h, Vh, r);
    }
#pragma unroll
    for (int i = 0; i < 6; i++) acc += (double)r[i] * (double)r[i];
    // xs is FP32
    float x[6];
    ld3<P32>((const char *)xs.data + off * sizeof(float2), nullptr, h, Vh, x);
#pragma unroll
    for (int i = 0; i < 6; i++) x[i] = fmaf((float)al, (float)p[i], x[i]);
    st3<P32>((char *)xs.data + off * sizeof(float2), nullptr, h, Vh, x);
  }
  const int nch = (Vh + 63) / 64;
  acc = bsum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) partC[(long)b * nch + (h >> 6)] = acc;
}

// new r2, reliable-update and stopping decisions, beta
__global__ void __launch_bounds__(256)
bs_scalar2(BsRhs *rhs, const int *active, BsGlobal *glob, const double *partC, int nch) {
  __shared__ double sh[4];
  if (glob->status != 0) return;
  const int b = blockIdx.x;
  BsRhs &R = rhs[b];
  if (!active[b]) return;
  double r2new = bs_block_sum(partC + (long)b * nch, nch, sh);
  if (threadIdx.x != 0) return;
  R.iter++;
  R.r2_old = R.r2;
  bool flag = false;
  if (!(R.pAp > 0)) { flag = true; }
  else if (r2new <= R.delta2 * R.r2_rel || r2new <= R.stop2 || R.iter >= R.max_iter) flag = true;
  R.r2 = r2new;
  if (flag) {
    R.need_update = 1;
    glob->pending = 1;  // becomes status at the end of the iteration
  } else {
    R.beta = r2new / R.r2_old;
  }
}

// ps = rs + beta ps
template <int P>
__global__ void __launch_bounds__(256)
bs_pupdate(const BsRhs *rhs, const int *active, const int *status, int Vh, VecView rs,
           VecView ps, long stride, int skip_flagged) {
  using R = typename BT<P>::R;
  if (status && *status != 0) return;
  const int b = blockIdx.y;
  if (!active[b]) return;
  if (skip_flagged && rhs[b].need_update) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= Vh) return;
  const R be = (R)rhs[b].beta;
  const size_t off = (size_t)b * stride;
  const long noff = (long)b * Vh;
  const size_t cb = sizeof(typename BT<P>::S);
  R r[6], p[6];
  ld3<P>((const char *)rs.data + off * cb, rs.norm ? rs.norm + noff : nullptr, h, Vh, r);
  ld3<P>((const char *)ps.data + off * cb, ps.norm ? ps.norm + noff : nullptr, h, Vh, p);
#pragma unroll
  for (int i = 0; i < 6; i++) p[i] = fma(be, p[i], r[i]);
  st3<P>((char *)ps.data + off * cb, ps.norm ? ps.norm + noff : nullptr, h, Vh, p);
}

// Last kernel of an iteration: raise the status the host polls.  Kernels of
// one iteration all see the same status because only this kernel changes it.
__global__ void bs_commit(BsGlobal *glob) {
  if (glob->status == 0 && glob->pending) glob->status = 1;
}

// out[b] = sum of slot b's partials (fixed order), for selected slots
__global__ void __launch_bounds__(256)
bs_sum(const double *part, int nch, const int *sel, double *out) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  if (sel && !sel[b]) return;
  double s = bs_block_sum(part + (long)b * nch, nch, sh);
  if (threadIdx.x == 0) out[b] = s;
}

void bs_launch_scalar1(BsRhs *rhs, const int *active, const int *status, const double *partA,
                       const double *partB, int nch, int n, hipStream_t st) {
  hipLaunchKernelGGL(bs_scalar1, dim3(n), dim3(256), 0, st, rhs, active, status, partA, partB, nch);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_update(BsRhs *rhs, const int *active, const int *status, int prec, int Vh,
                      VecView rs, VecView Aps, VecView ps, VecView xs, long stride,
                      double *partC, int n, hipStream_t st) {
  dim3 g((Vh + 255) / 256, n), bl(256);
  if (prec == P32) hipLaunchKernelGGL(bs_update<P32>, g, bl, 0, st, rhs, active, status, Vh, rs, Aps, ps, xs, stride, partC);
  else if (prec == P16) hipLaunchKernelGGL(bs_update<P16>, g, bl, 0, st, rhs, active, status, Vh, rs, Aps, ps, xs, stride, partC);
  else hipLaunchKernelGGL(bs_update<P64>, g, bl, 0, st, rhs, active, status, Vh, rs, Aps, ps, xs, stride, partC);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_scalar2(BsRhs *rhs, const int *active, BsGlobal *glob, const double *partC, int nch,
                       int n, hipStream_t st) {
  hipLaunchKernelGGL(bs_scalar2, dim3(n), dim3(256), 0, st, rhs, active, glob, partC, nch);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_commit(BsGlobal *glob, hipStream_t st) {
  hipLaunchKernelGGL(bs_commit, dim3(1), dim3(64), 0, st, glob);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_pupdate(const BsRhs *rhs, const int *active, const int *status, int prec, int Vh,
                       VecView rs, VecView ps, long stride, int n, int skip_flagged, hipStream_t st) {
  dim3 g((Vh + 255) / 256, n), bl(256);
  if (prec == P32) hipLaunchKernelGGL(bs_pupdate<P32>, g, bl, 0, st, rhs, active, status, Vh, rs, ps, stride, skip_flagged);
  else if (prec == P16) hipLaunchKernelGGL(bs_pupdate<P16>, g, bl, 0, st, rhs, active, status, Vh, rs, ps, stride, skip_flagged);
  else hipLaunchKernelGGL(bs_pupdate<P64>, g, bl, 0, st, rhs, active, status, Vh, rs, ps, stride, skip_flagged);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_sum(const double *part, int nch, int n, const int *sel, double *out, hipStream_t st) {
  hipLaunchKernelGGL(bs_sum, dim3(n), dim3(256), 0, st, part, nch, sel, out);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
