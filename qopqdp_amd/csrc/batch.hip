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

// Fixed-order sum of m partials by one 1024-thread block (four independent
// accumulators per thread keep several loads in flight).
constexpr int BS_THREADS = 1024;
__device__ double bs_block_sum(const double *p, int m, double *sh) {
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int i = threadIdx.x;
  for (; i + 3 * BS_THREADS < m; i += 4 * BS_THREADS) {
    s0 += p[i];
    s1 += p[i + BS_THREADS];
    s2 += p[i + 2 * BS_THREADS];
    s3 += p[i + 3 * BS_THREADS];
  }
  for (; i < m; i += BS_THREADS) s0 += p[i];
  double s = (s0 + s1) + (s2 + s3);
  s = bsum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = 0;
  if (threadIdx.x < 64) {
    r = threadIdx.x < BS_THREADS / 64 ? sh[threadIdx.x] : 0.0;
    r = bsum(r);
  }
  if (threadIdx.x == 0) sh[16] = r;
  __syncthreads();
  r = sh[16];
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

// Store v and replace it by the value as stored (FP16: rounded halves times
// the site norm, bit-identical to a reload).
template <int P>
__device__ __forceinline__ void st3_q(void *data, float *norm, long h, long Vh,
                                      typename BT<P>::R v[6]) {
  if constexpr (P == P16) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(v[i]));
    const float inv = m > 0.0f ? 1.0f / m : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      // each product and the packed half pair formed once (empty asm), so
      // the words read back are the words stored
      float a0 = v[2 * c] * inv, a1 = v[2 * c + 1] * inv;
      asm volatile("" : "+v"(a0), "+v"(a1));
      unsigned qw = __builtin_bit_cast(unsigned, __floats2half2_rn(a0, a1));
      asm volatile("" : "+v"(qw));
      ((unsigned *)data)[c * Vh + h] = qw;
      const float2 f = __half22float2(__builtin_bit_cast(__half2, qw));
      v[2 * c] = f.x * m;
      v[2 * c + 1] = f.y * m;
    }
    norm[h] = m;
  } else {
    st3<P>(data, norm, h, Vh, v);
  }
}

// pAp = |D p|^2 + m^2 |p|^2 ; alpha = r2 / pAp
__global__ void __launch_bounds__(1024)
bs_scalar1(BsRhs *rhs, const int *active, const int *status, const double *partA,
           const double *partB, int nch, int *ru_cont) {
  __shared__ double sh[17];
  if (*status != 0) return;
  const int b = blockIdx.x;
  BsRhs &R = rhs[b];
  if (threadIdx.x == 0) ru_cont[b] = 0;  // previous iteration's p fix is done
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
    st3_q<P>((char *)rs.data + off * cb, rs.norm ? rs.norm + noff : nullptr, h, Vh, r);
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
__global__ void __launch_bounds__(1024)
bs_scalar2(BsRhs *rhs, const int *active, BsGlobal *glob, const double *partC, int nch, int *ru) {
  __shared__ double sh[17];
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
    ru[b] = 1;  // reliable update later in this iteration
  } else {
    R.beta = r2new / R.r2_old;
  }
}

// ps = rs + beta ps
// ps = rs + beta ps ; partial |ps|^2 of the stored ps (the next iteration's
// p.Ap needs it)
template <int P>
__global__ void __launch_bounds__(256)
bs_pupdate(const BsRhs *rhs, const int *active, const int *status, int Vh, VecView rs,
           VecView ps, long stride, int skip_flagged, double *partB) {
  using R = typename BT<P>::R;
  if (status && *status != 0) return;
  const int b = blockIdx.y;
  if (!active[b]) return;
  if (skip_flagged && rhs[b].need_update) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  if (h < Vh) {
    const R be = (R)rhs[b].beta;
    const size_t off = (size_t)b * stride;
    const long noff = (long)b * Vh;
    const size_t cb = sizeof(typename BT<P>::S);
    R r[6], p[6];
    float *pn = ps.norm ? ps.norm + noff : nullptr;
    ld3<P>((const char *)rs.data + off * cb, rs.norm ? rs.norm + noff : nullptr, h, Vh, r);
    ld3<P>((const char *)ps.data + off * cb, pn, h, Vh, p);
#pragma unroll
    for (int i = 0; i < 6; i++) p[i] = fma(be, p[i], r[i]);
    st3_q<P>((char *)ps.data + off * cb, pn, h, Vh, p);
#pragma unroll
    for (int i = 0; i < 6; i++) acc += (double)p[i] * (double)p[i];
  }
  const int nch = (Vh + 63) / 64;
  acc = bsum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) partB[(long)b * nch + (h >> 6)] = acc;
}

// ---------------------------------------------------------------------------
// Reliable update on the device, for the slots flagged by bs_scalar2 (ru[b]):
//   x += xs, xs = 0                         (bs_ru_x)
//   ax = A x in FP64                        (two masked FP64 hops)
//   r = b - ax ; rs = r ; |r|^2, <rs, ps>   (bs_ru_resid)
//   stop / continue, beta                   (bs_ru_scalar)
//   ps = (1 - beta <rs,ps>/|r|^2) rs + beta ps  (bs_ru_pfix)
// Unflagged slots skip every step, so a slot's arithmetic is the same
// whatever the other slots do.  The vector steps run one block per 256 sites
// over the compacted list of the flagged slots (list[0] = their number, then
// the slots; bs_compact), so an iteration with no update costs one short
// launch per step.
__global__ void __launch_bounds__(256)
bs_ru_x(const int *list, int Vh, VecView x64, VecView xs, long stride) {
  const int k = list[0];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= Vh) return;
  for (int j = 0; j < k; j++) {
    const int b = list[1 + j];
    const size_t off = (size_t)b * stride;
    double2 *x = (double2 *)x64.data + off;
    float2 *y = (float2 *)xs.data + off;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 a = x[c * Vh + h];
      float2 d = y[c * Vh + h];
      a.x += (double)d.x;
      a.y += (double)d.y;
      x[c * Vh + h] = a;
      y[c * Vh + h] = make_float2(0.f, 0.f);
    }
  }
}

template <int P>
__global__ void __launch_bounds__(256)
bs_ru_resid(const int *list, int Vh, VecView b64, VecView ax64, VecView rs, VecView ps, long stride,
            double *part_r2, double *part_re, double *part_im) {
  using R = typename BT<P>::R;
  const int k = list[0];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = 0; j < k; j++) {
  const int b = list[1 + j];
  double r2 = 0, dre = 0, dim = 0;
  if (h < Vh) {
    const size_t off = (size_t)b * stride;
    const long noff = (long)b * Vh;
    const size_t cb = sizeof(typename BT<P>::S);
    const double2 *bb = (const double2 *)b64.data + off;
    const double2 *aa = (const double2 *)ax64.data + off;
    R rv[6];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double2 u = bb[c * Vh + h], w = aa[c * Vh + h];
      double xr = u.x - w.x, xi = u.y - w.y;
      r2 += xr * xr + xi * xi;
      rv[2 * c] = (R)xr;
      rv[2 * c + 1] = (R)xi;
    }
    float *rn = rs.norm ? rs.norm + noff : nullptr;
    st3_q<P>((char *)rs.data + off * cb, rn, h, Vh, rv);  // rv: the stored rs
    R p[6];
    const R *q = rv;
    ld3<P>((const char *)ps.data + off * cb, ps.norm ? ps.norm + noff : nullptr, h, Vh, p);
#pragma unroll
    for (int c = 0; c < 3; c++) {  // <rs, ps> = sum conj(rs) ps
      double qr = q[2 * c], qi = q[2 * c + 1], pr = p[2 * c], pi = p[2 * c + 1];
      dre += qr * pr + qi * pi;
      dim += qr * pi - qi * pr;
    }
  }
  const int nch = (Vh + 63) / 64;
  r2 = bsum(r2);
  dre = bsum(dre);
  dim = bsum(dim);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) {
    part_r2[(long)b * nch + (h >> 6)] = r2;
    part_re[(long)b * nch + (h >> 6)] = dre;
    part_im[(long)b * nch + (h >> 6)] = dim;
  }
  }
}

__global__ void __launch_bounds__(1024)
bs_ru_scalar(BsRhs *rhs, int *active, int *ru, int *ru_cont, double *lcoef,
             const double *part_r2, const double *part_re, const double *part_im, int nch) {
  __shared__ double sh[17];
  const int b = blockIdx.x;
  if (!ru[b]) return;
  double r2t = bs_block_sum(part_r2 + (long)b * nch, nch, sh);
  double rpr = bs_block_sum(part_re + (long)b * nch, nch, sh);
  double rpi = bs_block_sum(part_im + (long)b * nch, nch, sh);
  if (threadIdx.x != 0) return;
  BsRhs &R = rhs[b];
  R.nupdate++;
  R.need_update = 0;
  ru[b] = 0;
  const double r2_old = R.r2_old;
  R.r2 = r2t;
  R.r2_rel = r2t;
  if (r2t <= R.stop2) { R.reason = 1; active[b] = 0; return; }
  if (R.iter >= R.max_iter) { R.reason = 2; active[b] = 0; return; }
  if (!(R.pAp > 0) || !(r2t > 0)) { R.reason = 3; active[b] = 0; return; }
  const double beta = r2t / r2_old;
  R.beta = beta;
  lcoef[6 * b + 0] = 1.0 - beta * rpr / r2t;
  lcoef[6 * b + 1] = -beta * rpi / r2t;
  lcoef[6 * b + 2] = beta;
  lcoef[6 * b + 3] = 0.0;
  ru_cont[b] = 1;
}

template <int P>
__global__ void __launch_bounds__(256)
bs_ru_pfix(const int *list, const int *ru_cont, const double *lcoef, int Vh, VecView rs, VecView ps,
           long stride, double *partB) {
  using R = typename BT<P>::R;
  const int k = list[0];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = 0; j < k; j++) {
  const int b = list[1 + j];
  if (!ru_cont[b]) continue;  // the slot stopped at this update
  double acc = 0;
  if (h < Vh) {
  const size_t off = (size_t)b * stride;
  const long noff = (long)b * Vh;
  const size_t cb = sizeof(typename BT<P>::S);
  const double c0r = lcoef[6 * b], c0i = lcoef[6 * b + 1], c1 = lcoef[6 * b + 2];
  R r[6], p[6];
  ld3<P>((const char *)rs.data + off * cb, rs.norm ? rs.norm + noff : nullptr, h, Vh, r);
  ld3<P>((const char *)ps.data + off * cb, ps.norm ? ps.norm + noff : nullptr, h, Vh, p);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double rr = r[2 * c], ri = r[2 * c + 1];
    double pr = fma(c0r, rr, fma(-c0i, ri, c1 * (double)p[2 * c]));
    double pi = fma(c0r, ri, fma(c0i, rr, c1 * (double)p[2 * c + 1]));
    p[2 * c] = (R)pr;
    p[2 * c + 1] = (R)pi;
  }
  float *pn = ps.norm ? ps.norm + noff : nullptr;
  st3_q<P>((char *)ps.data + off * cb, pn, h, Vh, p);
#pragma unroll
  for (int i = 0; i < 6; i++) acc += (double)p[i] * (double)p[i];
  }
  const int nch = (Vh + 63) / 64;
  acc = bsum(acc);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nch) partB[(long)b * nch + (h >> 6)] = acc;
  }
}

// One thread: if any slot is flagged for a reliable update in this
// iteration, flag every running slot whose iterated residual has already
// dropped to join2 * r2_rel, so that the updates share the FP64 link reads.
// (A joined slot's update then depends on its batch-mates.)
__global__ void bs_ru_join(BsRhs *rhs, const int *active, int *ru, int n) {
  if (threadIdx.x != 0) return;
  bool any = false;
  for (int b = 0; b < n; b++) any |= active[b] && ru[b];
  if (!any) return;
  for (int b = 0; b < n; b++) {
    BsRhs &R = rhs[b];
    if (!active[b] || ru[b] || !(R.join2 > 0) || !(R.r2 <= R.join2 * R.r2_rel)) continue;
    R.need_update = 1;
    ru[b] = 1;
  }
}

// Last kernel of an iteration: raise the status the host polls.  Kernels of
// one iteration all see the same status because only this kernel changes it.
__global__ void bs_commit(BsGlobal *glob) {
  if (glob->status == 0 && glob->pending) glob->status = 1;
}

// out[b] = sum of slot b's partials (fixed order), for selected slots
__global__ void __launch_bounds__(1024)
bs_sum(const double *part, int nch, const int *sel, double *out) {
  __shared__ double sh[17];
  const int b = blockIdx.x;
  if (sel && !sel[b]) return;
  double s = bs_block_sum(part + (long)b * nch, nch, sh);
  if (threadIdx.x == 0) out[b] = s;
}

void bs_launch_scalar1(BsRhs *rhs, const int *active, const int *status, const double *partA,
                       const double *partB, int nch, int n, int *ru_cont, hipStream_t st) {
  hipLaunchKernelGGL(bs_scalar1, dim3(n), dim3(BS_THREADS), 0, st, rhs, active, status, partA, partB, nch, ru_cont);
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
                       int n, int *ru, hipStream_t st) {
  hipLaunchKernelGGL(bs_scalar2, dim3(n), dim3(BS_THREADS), 0, st, rhs, active, glob, partC, nch, ru);
  QOP_CHECK_HIP(hipGetLastError());
}
// Compacted list of the flagged slots: out[0] = their number k, out[1..k]
// their indices in increasing order (one wavefront, ballot and prefix count).
__global__ void __launch_bounds__(64) bs_compact(const int *flags, int n, int *out) {
  const int lane = (int)threadIdx.x;
  int base = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const bool f = i < n && flags[i] != 0;
    const unsigned long long m = __ballot(f);
    if (f) out[1 + base + __popcll(m & ((1ull << lane) - 1ull))] = i;
    base += __popcll(m);
  }
  if (lane == 0) out[0] = base;
}
void bs_launch_compact(const int *flags, int n, int *out, hipStream_t st) {
  hipLaunchKernelGGL(bs_compact, dim3(1), dim3(64), 0, st, flags, n, out);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_ru_join(BsRhs *rhs, const int *active, int *ru, int n, hipStream_t st) {
  hipLaunchKernelGGL(bs_ru_join, dim3(1), dim3(64), 0, st, rhs, active, ru, n);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_commit(BsGlobal *glob, hipStream_t st) {
  hipLaunchKernelGGL(bs_commit, dim3(1), dim3(64), 0, st, glob);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_pupdate(const BsRhs *rhs, const int *active, const int *status, int prec, int Vh,
                       VecView rs, VecView ps, long stride, int n, int skip_flagged, double *partB,
                       hipStream_t st) {
  dim3 g((Vh + 255) / 256, n), bl(256);
  if (prec == P32) hipLaunchKernelGGL(bs_pupdate<P32>, g, bl, 0, st, rhs, active, status, Vh, rs, ps, stride, skip_flagged, partB);
  else if (prec == P16) hipLaunchKernelGGL(bs_pupdate<P16>, g, bl, 0, st, rhs, active, status, Vh, rs, ps, stride, skip_flagged, partB);
  else hipLaunchKernelGGL(bs_pupdate<P64>, g, bl, 0, st, rhs, active, status, Vh, rs, ps, stride, skip_flagged, partB);
  QOP_CHECK_HIP(hipGetLastError());
}
void bs_launch_sum(const double *part, int nch, int n, const int *sel, double *out, hipStream_t st) {
  hipLaunchKernelGGL(bs_sum, dim3(n), dim3(BS_THREADS), 0, st, part, nch, sel, out);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop

namespace qop {
void bs_launch_ru(BsRhs *rhs, int *active, int *ru, int *ru_cont, const int *list, double *lcoef,
                  int prec, int Vh, VecView x64, VecView xs, VecView b64, VecView ax64, VecView rs,
                  VecView ps, long stride, double *pr2, double *pre, double *pim, int nch, int n,
                  int stage, double *ppB, hipStream_t st) {
  dim3 g((Vh + 255) / 256), bl(256);
  if (stage == 0) {
    hipLaunchKernelGGL(bs_ru_x, g, bl, 0, st, list, Vh, x64, xs, stride);
  } else if (stage == 1) {
    if (prec == P32) hipLaunchKernelGGL(bs_ru_resid<P32>, g, bl, 0, st, list, Vh, b64, ax64, rs, ps, stride, pr2, pre, pim);
    else if (prec == P16) hipLaunchKernelGGL(bs_ru_resid<P16>, g, bl, 0, st, list, Vh, b64, ax64, rs, ps, stride, pr2, pre, pim);
    else hipLaunchKernelGGL(bs_ru_resid<P64>, g, bl, 0, st, list, Vh, b64, ax64, rs, ps, stride, pr2, pre, pim);
  } else if (stage == 2) {
    hipLaunchKernelGGL(bs_ru_scalar, dim3(n), dim3(BS_THREADS), 0, st, rhs, active, ru, ru_cont, lcoef,
                       (const double *)pr2, (const double *)pre, (const double *)pim, nch);
  } else {
    if (prec == P32) hipLaunchKernelGGL(bs_ru_pfix<P32>, g, bl, 0, st, list, (const int *)ru_cont, (const double *)lcoef, Vh, rs, ps, stride, ppB);
    else if (prec == P16) hipLaunchKernelGGL(bs_ru_pfix<P16>, g, bl, 0, st, list, (const int *)ru_cont, (const double *)lcoef, Vh, rs, ps, stride, ppB);
    else hipLaunchKernelGGL(bs_ru_pfix<P64>, g, bl, 0, st, list, (const int *)ru_cont, (const double *)lcoef, Vh, rs, ps, stride, ppB);
  }
  QOP_CHECK_HIP(hipGetLastError());
}
}  // namespace qop
