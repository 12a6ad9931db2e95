// HIP kernels for the asqtad hot path on gfx950 (MI355X):
//  * the even/odd staggered hop (1-hop fat + 3-hop Naik, 16 double-stored
//    links per output site), single and multi right-hand side, in FP64,
//    FP32 and FP16 (+ per-link scale / per-site norm), with a fused
//    axpy epilogue and FP64 wavefront-shuffle reductions;
//  * FP64 reductions (norm2, complex dot) and elementwise vector kernels;
//  * raw <-> structure-of-arrays pack/unpack.
//
// The hop computes, for each output site x of parity p,
//   out(x) = sum_k W_k(x) psi(x + d_k)        (lib/asqtad_dslash_p.c:1246-1279)
// with d_k = (+mu, +3mu, -mu, -3mu) for k = 4mu+j (lib/asqtad_invert.c:27-36).
//
// Reductions are done per 64-site chunk (one wavefront) into a partial array
// indexed by chunk, then summed in a fixed order by launch_finalize, so each
// right-hand side's result does not depend on the block size, on the number of
// right-hand sides per thread or on which other right-hand sides share the
// launch.
#include "common.h"
#include <type_traits>
#include "kernels.h"

namespace qop {

// ---------------------------------------------------------------------------
// precision policies
template <int P> struct PT;
template <> struct PT<P64> { using R = double; using S = double2; };
template <> struct PT<P32> { using R = float;  using S = float2; };
template <> struct PT<P16> { using R = float;  using S = __half2; };

// Link element load.  NT: non-temporal (streaming) load, so the links, read
// once per hop, do not evict the gathered source vectors from L2 / MALL.
template <int P, bool NT>
__device__ __forceinline__ void ld_c(const void *base, long i, float scl,
                                     typename PT<P>::R &re, typename PT<P>::R &im) {
  if constexpr (P == P64) {
    const double *p = (const double *)base + 2 * i;
    if constexpr (NT) { re = __builtin_nontemporal_load(p); im = __builtin_nontemporal_load(p + 1); }
    else { double2 v = ((const double2 *)base)[i]; re = v.x; im = v.y; }
  } else if constexpr (P == P32) {
    if constexpr (NT) {
      const double *p = (const double *)base + i;  // one 8-byte load of the float2
      double d = __builtin_nontemporal_load(p);
      float2 v = __builtin_bit_cast(float2, d);
      re = v.x; im = v.y;
    } else { float2 v = ((const float2 *)base)[i]; re = v.x; im = v.y; }
  } else {
    unsigned u;
    if constexpr (NT) u = __builtin_nontemporal_load((const unsigned *)base + i);
    else u = ((const unsigned *)base)[i];
    __half2 v = __builtin_bit_cast(__half2, u);
    float2 f = __half22float2(v);
    re = f.x * scl; im = f.y * scl;
  }
}

// Stored words of one link (9 complex elements, + the per-link scale of FP16
// links), loaded one direction ahead of their use so that the loads stay in
// flight while the previous direction's products are formed.
template <int LP> struct LinkWords;
template <> struct LinkWords<P64> { double2 e[9]; };
template <> struct LinkWords<P32> { float2 e[9]; };
template <> struct LinkWords<P16> { unsigned e[9]; float s; };

template <int LP, bool NT>
__device__ __forceinline__ void link_words_load(LinkWords<LP> &w, const void *links,
                                                const float *lscale, long k, long Vh, long h) {
  const long base = k * 9 * Vh + h;
  if constexpr (LP == P16) {
    if constexpr (NT) w.s = __builtin_nontemporal_load(lscale + k * Vh + h);
    else w.s = lscale[k * Vh + h];
  }
#pragma unroll
  for (int e = 0; e < 9; e++) {
    const long i = base + (long)e * Vh;
    if constexpr (LP == P64) {
      const double *p = (const double *)links + 2 * i;
      if constexpr (NT) w.e[e] = make_double2(__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1));
      else w.e[e] = ((const double2 *)links)[i];
    } else if constexpr (LP == P32) {
      if constexpr (NT) w.e[e] = __builtin_bit_cast(float2, __builtin_nontemporal_load((const double *)links + i));
      else w.e[e] = ((const float2 *)links)[i];
    } else {
      if constexpr (NT) w.e[e] = __builtin_nontemporal_load((const unsigned *)links + i);
      else w.e[e] = ((const unsigned *)links)[i];
    }
  }
}

// The link as 18 reals: the same conversion as ld_c.
template <int LP, typename R>
__device__ __forceinline__ void link_words_cvt(const LinkWords<LP> &w, R out[18]) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    if constexpr (LP == P16) {
      // stored as (re, -im) / scale
      float2 f = __half22float2(__builtin_bit_cast(__half2, w.e[e]));
      out[2 * e] = f.x * w.s; out[2 * e + 1] = -(f.y * w.s);
    } else {
      out[2 * e] = w.e[e].x; out[2 * e + 1] = w.e[e].y;
    }
  }
}

// Buffer loads: a resource over the RHS group's vectors (or one direction's
// links), so a load takes one per-thread 32-bit byte offset register (site
// index x element size) and a uniform scalar offset (RHS, colour, link
// element) instead of a 64-bit address formed per load.  launch_hop_args
// uses them only where every offset stays below 2^31, the resource's range.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, 0x7fffffff, 0x00020000);
}

template <class T, bool NT>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  constexpr int aux = NT ? 2 : 0;  // 2: non-temporal
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, aux));
  } else {
    static_assert(sizeof(T) == 8, "buffer loads of 4 or 8 bytes");
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, aux));
  }
}

// link_words_load through buffer loads (FP32 / FP16 links): ho = h x element
// bytes, hs = 4 h.
template <int LP, bool NT>
__device__ __forceinline__ void link_words_load_buf(LinkWords<LP> &w, const void *links,
                                                    const float *lscale, long k, long Vh,
                                                    unsigned ho, unsigned hs) {
  static_assert(LP == P32 || LP == P16, "buffer link loads: FP32 or FP16 links");
  using E = typename std::conditional<LP == P32, float2, unsigned>::type;
  if constexpr (LP == P16) w.s = buf_load<float, NT>(buf_rsrc(lscale + k * Vh), hs, 0);
  const auto r = buf_rsrc((const E *)links + k * 9 * Vh);
#pragma unroll
  for (int e = 0; e < 9; e++) w.e[e] = buf_load<E, NT>(r, ho, (unsigned)(e * Vh * (long)sizeof(E)));
}

// Load the 3 colour components of site h from a parity-half view (NT:
// non-temporal loads, FP32 / FP16).
template <int P, bool NT = false>
__device__ __forceinline__ void vload(const void *data, const float *norm, long h,
                                      long Vh, typename PT<P>::R v[6]) {
  float n = 1.0f;
  if constexpr (P == P16) n = NT ? __builtin_nontemporal_load(norm + h) : norm[h];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if constexpr (P == P64) {
      double2 t = ((const double2 *)data)[c * Vh + h];
      v[2 * c] = t.x; v[2 * c + 1] = t.y;
    } else if constexpr (P == P32) {
      float2 t = NT ? __builtin_bit_cast(float2, __builtin_nontemporal_load((const double *)data + c * Vh + h))
                    : ((const float2 *)data)[c * Vh + h];
      v[2 * c] = t.x; v[2 * c + 1] = t.y;
    } else {
      const __half2 w = NT ? __builtin_bit_cast(__half2, __builtin_nontemporal_load((const unsigned *)data + c * Vh + h))
                           : ((const __half2 *)data)[c * Vh + h];
      float2 t = __half22float2(w);
      v[2 * c] = t.x * n; v[2 * c + 1] = t.y * n;
    }
  }
}

template <int P>
__device__ __forceinline__ void vstore(void *data, float *norm, long h, long Vh,
                                       const typename PT<P>::R v[6]) {
  if constexpr (P == P64) {
#pragma unroll
    for (int c = 0; c < 3; c++)
      ((double2 *)data)[c * Vh + h] = make_double2(v[2 * c], v[2 * c + 1]);
  } else if constexpr (P == P32) {
#pragma unroll
    for (int c = 0; c < 3; c++)
      ((float2 *)data)[c * Vh + h] = make_float2(v[2 * c], v[2 * c + 1]);
  } else {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(v[i]));
    float inv = m > 0.0f ? 1.0f / m : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++)
      ((__half2 *)data)[c * Vh + h] =
          __floats2half2_rn(v[2 * c] * inv, v[2 * c + 1] * inv);
    norm[h] = m;
  }
}

// Store v and replace it by the value as stored (for FP16 the rounded
// half-precision components times the site norm, bit-identical to loading it
// back), so reductions of the stored vector need no reload.
// NT: non-temporal (streaming) stores, so written vectors do not displace
// the gathered source vectors from L2.
template <int P, bool NT = false>
__device__ __forceinline__ void vstore_q(void *data, float *norm, long h, long Vh,
                                         typename PT<P>::R v[6]) {
  if constexpr (P == P16) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(v[i]));
    const float inv = m > 0.0f ? 1.0f / m : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      // each product and the packed half pair are formed once (empty asm),
      // so the words read back are the words stored (the compiler may
      // otherwise round v inv to FP16 in one step for one use and via FP32
      // for another)
      float a0 = v[2 * c] * inv, a1 = v[2 * c + 1] * inv;
      asm volatile("" : "+v"(a0), "+v"(a1));
      unsigned qw = __builtin_bit_cast(unsigned, __floats2half2_rn(a0, a1));
      asm volatile("" : "+v"(qw));
      if constexpr (NT) __builtin_nontemporal_store(qw, (unsigned *)data + c * Vh + h);
      else ((unsigned *)data)[c * Vh + h] = qw;
      const float2 f = __half22float2(__builtin_bit_cast(__half2, qw));
      v[2 * c] = f.x * m;
      v[2 * c + 1] = f.y * m;
    }
    if constexpr (NT) __builtin_nontemporal_store(m, norm + h);
    else norm[h] = m;
  } else if constexpr (NT && P == P32) {
#pragma unroll
    for (int c = 0; c < 3; c++)
      __builtin_nontemporal_store(__builtin_bit_cast(double, make_float2(v[2 * c], v[2 * c + 1])),
                                  (double *)data + c * Vh + h);
  } else if constexpr (NT && P == P64) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      __builtin_nontemporal_store(v[2 * c], (double *)data + 2 * (c * Vh + h));
      __builtin_nontemporal_store(v[2 * c + 1], (double *)data + 2 * (c * Vh + h) + 1);
    }
  } else {
    vstore<P>(data, norm, h, Vh, v);
  }
}

// Store v as vstore_q does and return |stored v|^2 in FP64.  FP16: the site
// norm m times the squared half words, m^2 sum(q^2), the sum over the 6 words
// in FP32 and the rest in FP64 (the stored value is q m).
template <int P, bool NT = false>
__device__ __forceinline__ double vstore_norm2(void *data, float *norm, long h, long Vh,
                                               typename PT<P>::R v[6]) {
  if constexpr (P == P16) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; i++) m = fmaxf(m, fabsf(v[i]));
    const float inv = m > 0.0f ? 1.0f / m : 0.0f;
    float q2 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      // each product and the packed half pair are formed once (empty asm):
      // the compiler may otherwise round v inv to FP16 directly (one
      // rounding) for one use and via FP32 (two) for another, so the words
      // squared would not always be the words stored
      float a0 = v[2 * c] * inv, a1 = v[2 * c + 1] * inv;
      asm volatile("" : "+v"(a0), "+v"(a1));
      unsigned qw = __builtin_bit_cast(unsigned, __floats2half2_rn(a0, a1));
      asm volatile("" : "+v"(qw));
      if constexpr (NT) __builtin_nontemporal_store(qw, (unsigned *)data + c * Vh + h);
      else ((unsigned *)data)[c * Vh + h] = qw;
      float2 f = __half22float2(__builtin_bit_cast(__half2, qw));
      asm volatile("" : "+v"(f.x), "+v"(f.y));
      q2 = fmaf(f.x, f.x, q2);
      q2 = fmaf(f.y, f.y, q2);
    }
    if constexpr (NT) __builtin_nontemporal_store(m, norm + h);
    else norm[h] = m;
    return (double)q2 * ((double)m * (double)m);
  } else {
    vstore_q<P, NT>(data, norm, h, Vh, v);
    double s = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) s += (double)v[i] * (double)v[i];
    return s;
  }
}

// ---------------------------------------------------------------------------
// wavefront reduction of a double (64 lanes) via cross-lane shuffles
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Site coordinates of output site h on parity p (raw order, x3 fastest).
__device__ __forceinline__ void site_coords(const Geom &g, int h, int p, int x[4],
                                            int &lex) {
  int t = h / g.Lh3;
  int x3h = h - t * g.Lh3;
  int t2 = t / g.L[2];
  x[2] = t - t2 * g.L[2];
  x[0] = t2 / g.L[1];
  x[1] = t2 - x[0] * g.L[1];
  x[3] = 2 * x3h + ((x[0] + x[1] + x[2] + p) & 1);
  lex = ((x[0] * g.L[1] + x[1]) * g.L[2] + x[2]) * g.L[3] + x[3];
}

__device__ __forceinline__ int nbr(const Geom &g, const int x[4], int lex, int mu,
                                   int d) {
  int y = x[mu] + d;
  int l = lex + d * g.stride[mu];
  if (y >= g.L[mu]) l -= g.L[mu] * g.stride[mu];
  else if (y < 0) l += g.L[mu] * g.stride[mu];
  return l >> 1;
}

// XCD-aware block remap: hardware dispatches block b to XCD b%8; give each
// XCD a contiguous range of logical blocks so neighbouring sites share an L2.
__device__ __forceinline__ int logical_block(int remap) {
  if (!remap) return blockIdx.x;
  int per = gridDim.x >> 3;
  return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// Output site of this thread for site block sb in the plane-blocked order
// (LDS staging mode 2): a block covers the same 64 positions of the x0 slab
// in blockDim/64 consecutive x0 planes, one plane per wave, so a wave's 64
// sites are still one 64-site reduction chunk.  Blocks past the last site
// return Vh (no site).
__device__ __forceinline__ int site_of_plane_block(const Geom &g, int sb) {
  const int D = blockDim.x >> 6;
  const int slab = g.Vh / g.L[0];
  const int ntile = slab >> 6;
  if (sb >= (g.L[0] / D) * ntile) return g.Vh;
  const int pg = sb / ntile, t = sb - pg * ntile;
  return (pg * D + (int)(threadIdx.x >> 6)) * slab + t * 64 + (int)(threadIdx.x & 63);
}

// Output site of this thread for site block sb.  Plain order: blocks cover
// consecutive sites.  x0-tiled order (tile > 0 sites, dividing the Vh/L0
// sites of an x0 slab): consecutive blocks sweep x0 for one tile of the slab
// before moving to the next tile, so the gathered x0 +-1, +-3 neighbours are
// reused from L2 a few blocks later instead of one slab later.  Blocks past
// the last site return Vh (no site).
__device__ __forceinline__ int site_of_block(const Geom &g, int sb, int tile) {
  if (!tile) return sb * blockDim.x + threadIdx.x;
  const int nblk = g.Vh / blockDim.x;
  if (sb >= nblk) return g.Vh;
  const int tb = tile / blockDim.x;
  const int col = g.L[0] * tb;
  const int j = sb / col;
  const int rem = sb - j * col;
  const int x0 = rem / tb;
  return x0 * (g.Vh / g.L[0]) + j * tile + (rem - x0 * tb) * blockDim.x + threadIdx.x;
}

// Output of one site for NB RHS: out = beta x + alpha acc (stored), or the
// fused CG update; per-RHS reduction terms in red[].
template <int P, int NB, int RED, bool HASX, bool NT>
__device__ __forceinline__ void hop_epilogue(const HopArgs &a, int h, int Vh, const int sl[NB],
                                             const bool act[NB], typename PT<P>::R acc[NB][6],
                                             double red[NB]) {
  using R = typename PT<P>::R;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if (!act[b]) continue;
    const int rb = sl[b];
    double al = a.alpha, be = a.beta;
    if (a.coef) { al = a.coef[2 * rb]; be = a.coef[2 * rb + 1]; }
    R o[6];
    R xv[6];
    if constexpr (HASX) {
      const char *xd = (const char *)a.x.data + (long)rb * a.xstride * sizeof(typename PT<P>::S);
      const float *xn = a.x.norm ? a.x.norm + (long)rb * a.nstride : nullptr;
      // the fused CG update streams p, r and x (each read or written once
      // per iteration at its own site): non-temporal
      vload<P, RED == RED_CG>(xd, xn, h, Vh, xv);
#pragma unroll
      for (int i = 0; i < 6; i++) o[i] = fma((R)be, xv[i], (R)al * acc[b][i]);
    } else {
#pragma unroll
      for (int i = 0; i < 6; i++) o[i] = (R)al * acc[b][i];
    }
    char *dd = (char *)a.dst.data + (long)rb * a.dstride * sizeof(typename PT<P>::S);
    float *dn = a.dst.norm ? a.dst.norm + (long)rb * a.nstride : nullptr;
    if constexpr (RED == RED_CG) {
      // fused CG update: o = A p (not stored); r -= alpha o ; x += alpha p
      // (x is FP32); reduction |r|^2 of the stored r
      const R alr = (R)a.cg_alpha[(long)rb * a.cg_stride];
      R rv[6];
      vload<P, true>(dd, dn, h, Vh, rv);
#pragma unroll
      for (int i = 0; i < 6; i++) rv[i] = fma(-alr, o[i], rv[i]);
      red[b] = vstore_norm2<P, true>(dd, dn, h, Vh, rv);
      // x is read again only at the next reliable update: non-temporal
      // load and store, so it does not displace the gathered neighbours in L2
      double *xs = (double *)a.xs.data + (long)rb * a.dstride;
      const float alf = (float)a.cg_alpha[(long)rb * a.cg_stride];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        float2 t = __builtin_bit_cast(float2, __builtin_nontemporal_load(xs + c * Vh + h));
        t.x = fmaf(alf, (float)xv[2 * c], t.x);
        t.y = fmaf(alf, (float)xv[2 * c + 1], t.y);
        __builtin_nontemporal_store(__builtin_bit_cast(double, t), xs + c * Vh + h);
      }
      continue;
    }
    if constexpr (RED == RED_OUT2) {
      red[b] = vstore_norm2<P, NT>(dd, dn, h, Vh, o);
      continue;
    }
    vstore_q<P, NT>(dd, dn, h, Vh, o);  // o becomes the stored value
    if constexpr (RED == RED_X2) {
      double s = 0;
#pragma unroll
      for (int i = 0; i < 6; i++) s += (double)xv[i] * (double)xv[i];
      red[b] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// hop kernel
// STG: LDS staging.  Each thread first loads the stored source words of the
// source-parity site with its own index into LDS; a gathered neighbour that
// is one of those sites is then read from LDS, the rest from memory.  Mode 1
// (a.stage = 1): the block's sites are contiguous, [h0, h0 + blockDim) -- the
// same (x0, x1, x2) rows as its outputs, so all x3 hops and the x2 hops that
// stay inside the block's rows come from LDS (an x0 tile is a whole number of
// blocks).  Mode 2: the plane-blocked order (site_of_plane_block), so the x0
// hops that stay inside the block's planes and the in-plane hops inside the
// wave's 64 positions come from LDS.  Staged blocks take up to 512 threads.
// LP: precision of the stored links (P, or P16 for FP32 vectors with FP16
// links and per-link scales).
// One RHS group of the hop for one site block: the group starting at launch
// RHS b0_in and site block sb_in, or (negative) the ones the grid gives.
template <int P, int NL, int NB, int RED, bool HASX, bool NT, bool SGL, bool STG, int LP, bool BUF>
__device__ __forceinline__ void hop_group(const HopArgs &a, int b0_in, int sb_in) {
  static_assert(!BUF || (P != P64 && !SGL), "buffer loads: FP32 / FP16, double-stored links");
  using R = typename PT<P>::R;
  const Geom g = a.g;
  const int Vh = g.Vh;
  const int lb = logical_block(a.remap);
  // rhs_inner: the RHS groups of one site block are consecutive logical
  // blocks, so they run together on one XCD and share its L2 (links and the
  // neighbour window); otherwise the groups are the grid's y dimension.
  const int sb = sb_in >= 0 ? sb_in : a.rhs_inner ? lb / a.ngroups : lb;
  const bool planes = STG && a.stage == 2;
  const int h = planes ? site_of_plane_block(g, sb) : site_of_block(g, sb, a.tile);
  const int b0 = b0_in >= 0 ? b0_in : (a.rhs_inner ? lb % a.ngroups : (int)blockIdx.y) * NB;
  if (a.status && *a.status != 0) return;

  double red[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) red[b] = 0.0;

  bool act[NB];
  int sl[NB];  // slot of each RHS of this thread
  bool any = false;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    // the launch's RHS j = b0 + b is slot j, or with a compacted list
    // (a.slots: count, then the slots) the j-th listed slot
    const int j = b0 + b;
    if (a.slots) {
      act[b] = j < a.slots[0];
      sl[b] = act[b] ? a.slots[1 + j] : j;
    } else {
      act[b] = (j < a.n) && (!a.active || a.active[j]);
      sl[b] = j;
    }
    any |= act[b];
  }
  using S = typename PT<P>::S;
  const int bs = blockDim.x;
  const int h0 = h - (int)threadIdx.x;  // mode 1: first site of the block
  const int wv = (int)(threadIdx.x >> 6), nwv = bs >> 6;  // mode 2: wave = plane
  const int hrow = h - (int)(threadIdx.x & 63);         // mode 2: first site of the wave
  // LDS tile: FP16 one 16-byte record per RHS and site, [NB][bs] x {3
  // half2 words, norm} (one ds_read_b128 per gather, consecutive lanes 16 B
  // apart); FP32 [NB][3][bs] float2
  S *td = nullptr;
  float *tn = nullptr;
  uint4 *tr = nullptr;
  if constexpr (STG) {
    extern __shared__ __align__(16) unsigned char stg_mem[];
    td = (S *)stg_mem;
    tn = (float *)(td + NB * 3 * bs);
    tr = (uint4 *)stg_mem;
    const int hs = h;
    if (any && hs < Vh) {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (!act[b]) continue;
        const S *sd = (const S *)a.src.data + (long)sl[b] * a.sstride;
        if constexpr (P == P16) {
          uint4 rec;
          rec.x = __builtin_bit_cast(unsigned, sd[hs]);
          rec.y = __builtin_bit_cast(unsigned, sd[(long)Vh + hs]);
          rec.z = __builtin_bit_cast(unsigned, sd[2 * (long)Vh + hs]);
          rec.w = __builtin_bit_cast(unsigned, a.src.norm[(long)sl[b] * a.nstride + hs]);
          tr[b * bs + threadIdx.x] = rec;
        } else {
#pragma unroll
          for (int c = 0; c < 3; c++) td[(b * 3 + c) * bs + threadIdx.x] = sd[(long)c * Vh + hs];
        }
      }
    }
    __syncthreads();
  }
  // a group with no active slot reads nothing (masked launches are cheap)
  if (h < Vh && any) {
    int x[4], lex;
    site_coords(g, h, a.parity, x, lex);
    R acc[NB][6];
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int i = 0; i < 6; i++) acc[b][i] = 0;

    constexpr int PER = NL / 4;  // links per direction: 4 (fat+long) or 2
    // Per link: the NB gathers are issued first, then the next link's words,
    // and only then are the values converted and used: one wait per link
    // instead of one per RHS.  With one RHS per thread, and in FP64, each
    // link is loaded where it is used instead (there the prefetch costs more
    // registers than it hides).
    constexpr bool PF = NB > 1 && !SGL && P != P64;
    // BUF: gathers and link loads through buffer loads
    const auto rs_src = buf_rsrc((const S *)a.src.data + (long)b0 * a.sstride);
    const auto rs_nrm = buf_rsrc(P == P16 ? a.src.norm + (long)b0 * a.nstride : a.src.norm);
    const unsigned lho = (unsigned)h * (unsigned)(LP == P32 ? 8 : 4);
    auto load_link = [&](LinkWords<LP> &lw_, int kk) {
      if constexpr (BUF) link_words_load_buf<LP, NT>(lw_, a.links, a.lscale, kk, Vh, lho, (unsigned)h * 4u);
      else link_words_load<LP, NT>(lw_, a.links, a.lscale, kk, Vh, h);
    };
    LinkWords<LP> lnext;
    if constexpr (PF) load_link(lnext, 0);
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
#pragma unroll
      for (int j = 0; j < PER; j++) {
        const int d = PER == 4 ? (j == 0 ? 1 : j == 1 ? 3 : j == 2 ? -1 : -3)
                               : (j == 0 ? 1 : -1);
        const int k = PER * mu + j;
        const int nb = nbr(g, x, lex, mu, d);
        int slot = -1;  // LDS position of the neighbour, or -1
        if constexpr (STG) {
          if (planes) {
            if (mu == 0) {
              if ((unsigned)(wv + d) < (unsigned)nwv) slot = (int)threadIdx.x + 64 * d;
            } else {
              const unsigned pos = (unsigned)(nb - hrow);
              if (pos < 64u) slot = (wv << 6) + (int)pos;
            }
          } else {
            const unsigned o = (unsigned)(nb - h0);
            if (o < (unsigned)bs) slot = (int)o;
          }
        }
        const bool in_lds = STG && slot >= 0;
        LinkWords<LP> lcur;  // without the prefetch: this link, loaded before the gathers
        if constexpr (!PF && !SGL) load_link(lcur, k);
        // stored words of the neighbour for each RHS: a wave with any LDS
        // lane first reads the tile (lanes without a staged neighbour read
        // position 0), then the memory lanes load over their values, so the
        // memory loads wait only for the short LDS reads
        S gm[NB][3];
        float nm[NB];
        if constexpr (STG) {
          if (__any(in_lds)) {
            const int o = in_lds ? slot : 0;
#pragma unroll
            for (int b = 0; b < NB; b++) {
              if (!act[b]) continue;
              if constexpr (P == P16) {
                // one 16-byte read per record: the words pass through empty
                // asm as 32-bit integers, so the compiler does not split the
                // read by the types they are used as
                const uint4 rec = tr[b * bs + o];
                unsigned w0 = rec.x, w1 = rec.y, w2 = rec.z, w3 = rec.w;
                asm volatile("" : "+v"(w0), "+v"(w1), "+v"(w2), "+v"(w3));
                gm[b][0] = __builtin_bit_cast(S, w0);
                gm[b][1] = __builtin_bit_cast(S, w1);
                gm[b][2] = __builtin_bit_cast(S, w2);
                nm[b] = __builtin_bit_cast(float, w3);
              } else {
#pragma unroll
                for (int c = 0; c < 3; c++) gm[b][c] = td[(b * 3 + c) * bs + o];
              }
            }
          }
        }
        if (!in_lds) {
          const unsigned vo = (unsigned)nb * (unsigned)sizeof(S);
#pragma unroll
          for (int b = 0; b < NB; b++) {
            if (!act[b]) continue;
            if constexpr (BUF) {
#pragma unroll
              for (int c = 0; c < 3; c++)
                gm[b][c] = buf_load<S, false>(rs_src, vo, (unsigned)(((long)b * a.sstride + (long)c * Vh) * (long)sizeof(S)));
              if constexpr (P == P16)
                nm[b] = buf_load<float, false>(rs_nrm, (unsigned)nb * 4u, (unsigned)((long)b * a.nstride * 4));
            } else {
              const S *sd = (const S *)a.src.data + (long)sl[b] * a.sstride;
#pragma unroll
              for (int c = 0; c < 3; c++) gm[b][c] = sd[(long)c * Vh + nb];
              if constexpr (P == P16) nm[b] = a.src.norm[(long)sl[b] * a.nstride + nb];
            }
          }
        }
        R w[18];
        LinkWords<LP> lw;  // this link's stored words
        if constexpr (SGL) {
          // single-stored: forward link of this site, or -U(x-d)^dagger from
          // the forward link of the neighbour (source parity)
          constexpr int HP = PER / 2;
          const bool fwd = j < HP;  // resolved at compile time once j is unrolled
          const int ks = HP * mu + (fwd ? j : j - HP);
          const long lb2 = (long)ks * 9 * Vh + (fwd ? h : nb);
          const void *base = fwd ? a.links : a.links_bck;
          R u[18];
#pragma unroll
          for (int e = 0; e < 9; e++) ld_c<P, false>(base, lb2 + (long)e * Vh, 1.0f, u[2 * e], u[2 * e + 1]);
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
              if (fwd) {
                w[6 * r + 2 * c] = u[6 * r + 2 * c];
                w[6 * r + 2 * c + 1] = u[6 * r + 2 * c + 1];
              } else {
                w[6 * r + 2 * c] = -u[6 * c + 2 * r];
                w[6 * r + 2 * c + 1] = u[6 * c + 2 * r + 1];
              }
            }
        } else {
          if constexpr (PF) {
            lw = lnext;
            if (j + 1 < PER || mu < 3) load_link(lnext, k + 1);
          } else {
            lw = lcur;
          }
          if constexpr (P != P16) link_words_cvt<LP>(lw, w);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
          if (!act[b]) continue;
          if constexpr (P == P16) {
            // FP16 links and vectors: the products of the stored half-precision
            // words (no conversions), summed per colour row in FP32, then
            // scaled once by link scale x site norm.  The link word holds
            // (wr, -wi), so the real part wr vr - wi vi is one two-term dot
            // product (v_dot2c_f32_f16; FP16 subnormal inputs count as zero,
            // below the format's rounding at the link and site scales), the
            // imaginary part two mixed-precision FMAs with exact products.
            const float sn = lw.s * nm[b];
            float t[6];
#pragma unroll
            for (int r = 0; r < 3; r++) {
              float re = 0.0f, im = 0.0f;
#pragma unroll
              for (int c = 0; c < 3; c++) {
                using h2 = _Float16 __attribute__((ext_vector_type(2)));
                const h2 wd = __builtin_bit_cast(h2, lw.e[3 * r + c]);
                const h2 vd = __builtin_bit_cast(h2, gm[b][c]);
                re = __builtin_amdgcn_fdot2(wd, vd, re, false);
                const float wr = (float)wd.x, wni = (float)wd.y;  // wni = -wi
                const float vr = (float)vd.x, vi = (float)vd.y;
                im = fmaf(wr, vi, im);
                im = fmaf(-wni, vr, im);
              }
              t[2 * r] = re; t[2 * r + 1] = im;
            }
#pragma unroll
            for (int q = 0; q < 6; q++) acc[b][q] = fmaf(sn, t[q], acc[b][q]);
          } else {
            R v[6];
#pragma unroll
            for (int c = 0; c < 3; c++) {
              const S t = gm[b][c];
              v[2 * c] = t.x; v[2 * c + 1] = t.y;
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
              R re = acc[b][2 * r], im = acc[b][2 * r + 1];
#pragma unroll
              for (int c = 0; c < 3; c++) {
                const R wr = w[6 * r + 2 * c], wi = w[6 * r + 2 * c + 1];
                re = fma(wr, v[2 * c], re);
                re = fma(-wi, v[2 * c + 1], re);
                im = fma(wr, v[2 * c + 1], im);
                im = fma(wi, v[2 * c], im);
              }
              acc[b][2 * r] = re; acc[b][2 * r + 1] = im;
            }
          }
        }
      }
    }
    hop_epilogue<P, NB, RED, HASX, NT>(a, h, Vh, sl, act, acc, red);
  }
  if constexpr (RED != RED_NONE) {
    const long nchunk = (Vh + 63) / 64;
    const int chunk = h >> 6;
#pragma unroll
    for (int b = 0; b < NB; b++) {
      double s = wave_sum(red[b]);
      if ((threadIdx.x & 63) == 0 && chunk < nchunk && b0 + b < a.n)
        a.partial[(long)(b0 + b) * nchunk + chunk] = s;
    }
  }
}

// GL (FP64 launches with a compacted slot list): the RHS group is the
// innermost logical block index (XCD-aware order), so the groups of one site
// block run side by side and share its links in L2, and the blocks of groups
// past the listed slots return before any load: an empty list costs a
// launch of idle blocks.  Otherwise one group and site block per block.
template <int P, int NL, int NB, int RED, bool HASX, bool NT, bool SGL, bool STG, int LP = P,
          bool BUF = false, bool GL = false>
__global__ void __launch_bounds__(STG ? 512 : 256)
hop_kernel(const HopArgs a) {
  if constexpr (GL) {
    const int lb = logical_block(a.remap);
    const int b0 = (lb % a.ngroups) * NB, sb = lb / a.ngroups;
    if (b0 >= a.slots[0] || sb * (int)blockDim.x >= a.g.Vh) return;
    hop_group<P, NL, NB, RED, HASX, NT, SGL, STG, LP, BUF>(a, b0, sb);
  } else {
    hop_group<P, NL, NB, RED, HASX, NT, SGL, STG, LP, BUF>(a, -1, -1);
  }
}

int hop_partial_blocks(const Lattice *lat) { return (int)((lat->Vh + 63) / 64); }

// LDS bytes of a staged block: NB RHS x blockDim sites x (3 stored complex
// elements + the FP16 norm)
static size_t stage_bytes(int prec, int nb, int block) {
  const size_t el = prec == P64 ? 16 : 8;  // double2 or float2
  return prec == P16 ? (size_t)nb * block * (3 * 4 + 4) : (size_t)nb * block * 3 * el;
}

// BUF: buffer loads (FP32 / FP16 with double-stored links; launch_hop_args
// routes lattices whose offsets do not fit to the plain-load instance).
template <int P, int NL, int NB, bool NT, bool SGL = false, bool STG = false, int LP = P,
          bool BUF = (P != P64 && !SGL), bool GL = false>
static void hop_dispatch_red(const HopArgs &a, dim3 grid, dim3 block, int red,
                             bool hasx, hipStream_t st) {
  const size_t lds = STG ? stage_bytes(P, NB, (int)block.x) : 0;
  if constexpr (GL) {
    // compacted slot lists: FP64, no reduction
    if (hasx) hipLaunchKernelGGL((hop_kernel<P, NL, NB, RED_NONE, true, NT, SGL, STG, LP, BUF, true>), grid, block, lds, st, a);
    else hipLaunchKernelGGL((hop_kernel<P, NL, NB, RED_NONE, false, NT, SGL, STG, LP, BUF, true>), grid, block, lds, st, a);
    (void)red;
    return;
  }
#define QOP_LAUNCH(RD, HX) \
  hipLaunchKernelGGL((hop_kernel<P, NL, NB, RD, HX, NT, SGL, STG, LP, BUF>), grid, block, lds, st, a)
  if (red == RED_NONE) { if (hasx) QOP_LAUNCH(RED_NONE, true); else QOP_LAUNCH(RED_NONE, false); }
  else if (red == RED_OUT2) { if (hasx) QOP_LAUNCH(RED_OUT2, true); else QOP_LAUNCH(RED_OUT2, false); }
  else if (red == RED_CG) { QOP_LAUNCH(RED_CG, true); }
  else { QOP_LAUNCH(RED_X2, true); }
#undef QOP_LAUNCH
}

template <int P, int NL, int LP = P>
static void hop_dispatch_nb(const HopArgs &a, int nb, dim3 grid, dim3 block,
                            int red, bool hasx, hipStream_t st) {
  if (a.single) {
    if constexpr (P != P16 && LP == P) {
      QOP_REQUIRE(nb == 1 && !a.nt, "hop: single-stored links take one RHS per thread");
      hop_dispatch_red<P, NL, 1, false, true, false, LP>(a, grid, block, red, hasx, st);
    } else {
      fatal(__FILE__, __LINE__, "hop: single-stored links exist in FP64 and FP32 only");
    }
    return;
  }
  if (a.slots) {
    if constexpr (P == P64) {
      if (a.stage) {
        if constexpr (NL == 16) {
          if (nb == 2) hop_dispatch_red<P, NL, 2, false, false, true, LP, false, true>(a, grid, block, red, hasx, st);
          else hop_dispatch_red<P, NL, 4, false, false, true, LP, false, true>(a, grid, block, red, hasx, st);
        } else {
          fatal(__FILE__, __LINE__, "hop: LDS staging needs Naik links");
        }
        return;
      }
      switch (nb) {
        case 1: hop_dispatch_red<P, NL, 1, false, false, false, LP, false, true>(a, grid, block, red, hasx, st); break;
        case 2: hop_dispatch_red<P, NL, 2, false, false, false, LP, false, true>(a, grid, block, red, hasx, st); break;
        case 4: hop_dispatch_red<P, NL, 4, false, false, false, LP, false, true>(a, grid, block, red, hasx, st); break;
        default: hop_dispatch_red<P, NL, 8, false, false, false, LP, false, true>(a, grid, block, red, hasx, st); break;
      }
      return;
    }
  }
  if (!a.buf) {
    // offsets beyond the buffer range: one RHS per thread, plain loads
    if constexpr (P != P64) {
      QOP_REQUIRE(nb == 1 && !a.stage && !a.nt, "hop: plain-load instance takes one RHS per thread");
      hop_dispatch_red<P, NL, 1, false, false, false, LP, false>(a, grid, block, red, hasx, st);
      return;
    }
  }
  if (a.stage) {
    if constexpr (P == P64 && NL == 16) {
      // FP64: 2 or 4 RHS per thread (the tile of 8 exceeds 64 KiB at 256 threads)
      if (nb == 2) hop_dispatch_red<P, NL, 2, false, false, true, LP>(a, grid, block, red, hasx, st);
      else hop_dispatch_red<P, NL, 4, false, false, true, LP>(a, grid, block, red, hasx, st);
    } else if constexpr (NL == 16) {
      switch (nb) {
        case 1: hop_dispatch_red<P, NL, 1, false, false, true, LP>(a, grid, block, red, hasx, st); break;
        case 2: hop_dispatch_red<P, NL, 2, false, false, true, LP>(a, grid, block, red, hasx, st); break;
        case 4: hop_dispatch_red<P, NL, 4, false, false, true, LP>(a, grid, block, red, hasx, st); break;
        default: hop_dispatch_red<P, NL, 8, false, false, true, LP>(a, grid, block, red, hasx, st); break;
      }
    } else {
      fatal(__FILE__, __LINE__, "hop: LDS staging exists for FP32/FP16 with Naik links");
    }
    return;
  }
  if (a.nt) {
    switch (nb) {
      case 1: hop_dispatch_red<P, NL, 1, true, false, false, LP>(a, grid, block, red, hasx, st); break;
      case 2: hop_dispatch_red<P, NL, 2, true, false, false, LP>(a, grid, block, red, hasx, st); break;
      case 4: hop_dispatch_red<P, NL, 4, true, false, false, LP>(a, grid, block, red, hasx, st); break;
      case 8: hop_dispatch_red<P, NL, 8, true, false, false, LP>(a, grid, block, red, hasx, st); break;
      default:
        if constexpr (P != P64 && NL == 16) hop_dispatch_red<P, NL, 16, true, false, false, LP>(a, grid, block, red, hasx, st);
        else fatal(__FILE__, __LINE__, "hop: 16 RHS per thread needs FP32/FP16 and Naik links");
        break;
    }
  } else {
    switch (nb) {
      case 1: hop_dispatch_red<P, NL, 1, false, false, false, LP>(a, grid, block, red, hasx, st); break;
      case 2: hop_dispatch_red<P, NL, 2, false, false, false, LP>(a, grid, block, red, hasx, st); break;
      case 4: hop_dispatch_red<P, NL, 4, false, false, false, LP>(a, grid, block, red, hasx, st); break;
      case 8: hop_dispatch_red<P, NL, 8, false, false, false, LP>(a, grid, block, red, hasx, st); break;
      default:
        if constexpr (P != P64 && NL == 16) hop_dispatch_red<P, NL, 16, false, false, false, LP>(a, grid, block, red, hasx, st);
        else fatal(__FILE__, __LINE__, "hop: 16 RHS per thread needs FP32/FP16 and Naik links");
        break;
    }
  }
}

void launch_hop_args(HopArgs a, int prec, int nl, int nb, int block, int red,
                     bool hasx, hipStream_t st) {
  QOP_REQUIRE(nb == 1 || nb == 2 || nb == 4 || nb == 8 || nb == 16, "hop: rhs per thread must be 1,2,4,8,16");
  QOP_REQUIRE(block % 64 == 0 && block >= 64 && block <= (a.stage ? 512 : 256),
              "hop: block must be a multiple of 64, 64..256 (512 with LDS staging)");
  QOP_REQUIRE(red == RED_NONE || hasx || red == RED_OUT2, "hop: RED_X2 / RED_CG need x");
  QOP_REQUIRE(red != RED_CG || (a.cg_alpha && a.xs.data && prec != P64), "hop: RED_CG needs alpha and xs");
  QOP_REQUIRE(a.n >= 1, "hop: n >= 1");
  QOP_REQUIRE(!a.slots || (prec == P64 && !a.single && red == RED_NONE && nb <= 8),
              "hop: a compacted slot list is for FP64 double-stored hops without reductions, at "
              "most 8 RHS per thread");
  if (a.stage) {
    QOP_REQUIRE(nl == 16 && nb <= 8 && !a.single && !a.nt && (prec != P64 || nb == 2 || nb == 4),
                "hop: LDS staging needs Naik links, at most 8 RHS per thread (FP64: 2 or 4) and "
                "plain link loads");
    QOP_REQUIRE(stage_bytes(prec, nb, block) <= 65536, "hop: LDS staging tile above 64 KiB");
    QOP_REQUIRE(a.stage == 1 || a.stage == 2, "hop: LDS staging mode must be 1 (rows) or 2 (planes)");
    if (a.stage == 2) {
      const int slab = a.g.Vh / a.g.L[0];
      QOP_REQUIRE(!a.tile && slab % 64 == 0 && a.g.L[0] % (block / 64) == 0,
                  "hop: plane-blocked staging needs the plain order, 64 dividing the x0 slab and "
                  "block/64 dividing L0");
    }
  }
  if (a.tile) {
    const int slab = a.g.Vh / a.g.L[0];
    QOP_REQUIRE(a.tile % block == 0 && slab % a.tile == 0, "hop: tile must divide the x0 slab");
  }
  // buffer loads (FP32 / FP16, double-stored links) where every byte offset
  // from the RHS group's first vector / norm and from a direction's first
  // link stays below 2^31; otherwise (or when a.plain asks for it) one RHS
  // per thread with plain loads
  {
    const long lim = 1L << 31, Vh = a.g.Vh;
    const long esz = prec == P16 ? 4 : 8;   // stored vector element
    const long lesz = prec == P32 ? 8 : 4;  // stored link element (P32H: FP16 links)
    a.buf = prec != P64 && !a.single && !a.plain && ((long)(nb - 1) * a.sstride + 3 * Vh) * esz < lim &&
            ((long)(nb - 1) * a.nstride + Vh) * 4 < lim && 9 * Vh * lesz < lim;
    if (prec != P64 && !a.single && !a.buf) {
      nb = 1;
      a.stage = 0;
      a.nt = 0;
      if (block > 256) block = 256;
    }
  }
  int nblk = (a.g.Vh + block - 1) / block;
  a.ngroups = (a.n + nb - 1) / nb;
  if (a.slots) a.rhs_inner = 1;  // the GL kernel orders the groups innermost itself
  dim3 grid, blk(block);
  if (a.rhs_inner) {
    int tot = nblk * a.ngroups;
    if (a.remap) tot = (tot + 7) / 8 * 8;
    grid = dim3(tot, 1);
  } else {
    if (a.remap) nblk = (nblk + 7) / 8 * 8;
    grid = dim3(nblk, a.ngroups);
  }
  if (prec == P64) {
    if (nl == 16) hop_dispatch_nb<P64, 16>(a, nb, grid, blk, red, hasx, st);
    else hop_dispatch_nb<P64, 8>(a, nb, grid, blk, red, hasx, st);
  } else if (prec == P32H) {
    QOP_REQUIRE(!a.single, "hop: FP32 vectors with FP16 links take double-stored links");
    if (nl == 16) hop_dispatch_nb<P32, 16, P16>(a, nb, grid, blk, red, hasx, st);
    else hop_dispatch_nb<P32, 8, P16>(a, nb, grid, blk, red, hasx, st);
  } else if (prec == P32) {
    if (nl == 16) hop_dispatch_nb<P32, 16>(a, nb, grid, blk, red, hasx, st);
    else hop_dispatch_nb<P32, 8>(a, nb, grid, blk, red, hasx, st);
  } else {
    if (nl == 16) hop_dispatch_nb<P16, 16>(a, nb, grid, blk, red, hasx, st);
    else hop_dispatch_nb<P16, 8>(a, nb, grid, blk, red, hasx, st);
  }
  QOP_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// finalize: result[b] = sum_i partial[b*m + i], fixed order
__global__ void __launch_bounds__(256)
finalize_kernel(const double *partial, int m, int nval, double *result,
                const int *status) {
  if (status && *status != 0) return;
  __shared__ double sh[4];
  const int b = blockIdx.x;
  const double *p = partial + (long)b * m;
  double s = 0;
  for (int i = threadIdx.x; i < m; i += 256) s += p[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) result[b] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  (void)nval;
}

void launch_finalize(const double *partial, int m, int n, double *result,
                     hipStream_t st, const int *status) {
  hipLaunchKernelGGL(finalize_kernel, dim3(n), dim3(256), 0, st, partial, m, n,
                     result, status);
  QOP_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// site-wise vector kernels (one thread per site of a parity half, n slots)
//
// generic linear combination:
//   out = c0*v0 + c1*v1 (+ c2*v2), complex coefficients (host values or a
//   device array coef[slot][3][2]), with an optional reduction of |out|^2.
template <int PO, int P0, int P1, int P2, int NIN, bool RED>
__global__ void __launch_bounds__(256)
lincomb_kernel(const LincombArgs a) {
  const int Vh = a.Vh;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (a.status && *a.status != 0) return;
  if (a.active && !a.active[b]) return;  // inactive slot: no output, no partial
  double s = 0;
  if (h < Vh) {
    double cr[3], ci[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (a.coef) {
        cr[i] = a.coef[(long)b * 6 + 2 * i]; ci[i] = a.coef[(long)b * 6 + 2 * i + 1];
        if (a.coef_neg & (1 << i)) { cr[i] = -cr[i]; ci[i] = -ci[i]; }
      } else { cr[i] = a.cr[i]; ci[i] = a.ci[i]; }
    }
    double o[6] = {0, 0, 0, 0, 0, 0};
    auto accum = [&](auto tag, const VecView &v, long stride, double c_r, double c_i) {
      constexpr int PP = decltype(tag)::value;
      typename PT<PP>::R t[6];
      const char *d = (const char *)v.data + (long)b * stride * sizeof(typename PT<PP>::S);
      const float *nm = v.norm ? v.norm + (long)b * a.nstride : nullptr;
      vload<PP>(d, nm, h, Vh, t);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double tr = t[2 * c], ti = t[2 * c + 1];
        o[2 * c] = fma(c_r, tr, fma(-c_i, ti, o[2 * c]));
        o[2 * c + 1] = fma(c_r, ti, fma(c_i, tr, o[2 * c + 1]));
      }
    };
    accum(std::integral_constant<int, P0>{}, a.v[0], a.stride[0], cr[0], ci[0]);
    if constexpr (NIN > 1) accum(std::integral_constant<int, P1>{}, a.v[1], a.stride[1], cr[1], ci[1]);
    if constexpr (NIN > 2) accum(std::integral_constant<int, P2>{}, a.v[2], a.stride[2], cr[2], ci[2]);
    typename PT<PO>::R oo[6];
#pragma unroll
    for (int i = 0; i < 6; i++) oo[i] = (typename PT<PO>::R)o[i];
    char *d = (char *)a.out.data + (long)b * a.ostride * sizeof(typename PT<PO>::S);
    float *nm = a.out.norm ? a.out.norm + (long)b * a.nstride : nullptr;
    vstore_q<PO>(d, nm, h, Vh, oo);
    if constexpr (RED) {
#pragma unroll
      for (int i = 0; i < 6; i++) s += (double)oo[i] * (double)oo[i];
    }
  }
  if constexpr (RED) {
    const long nchunk = (Vh + 63) / 64;
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && (h >> 6) < nchunk) a.partial[(long)b * nchunk + (h >> 6)] = s;
  }
}

template <int PO, int P0, int P1, int P2>
static void lincomb_dispatch2(const LincombArgs &a, int nin, bool red, dim3 g,
                              dim3 bl, hipStream_t st) {
#define QOP_L(N, RD) hipLaunchKernelGGL((lincomb_kernel<PO, P0, P1, P2, N, RD>), g, bl, 0, st, a)
  if (nin == 1) { if (red) QOP_L(1, true); else QOP_L(1, false); }
  else if (nin == 2) { if (red) QOP_L(2, true); else QOP_L(2, false); }
  else { if (red) QOP_L(3, true); else QOP_L(3, false); }
#undef QOP_L
}

// Supported precision combinations: all same precision, or mixed
// (out FP64 with FP32/FP16 inputs for accumulations, and conversions).
void launch_lincomb(const LincombArgs &a, int po, const int *pin, int nin,
                    bool red, hipStream_t st) {
  dim3 g((a.Vh + 255) / 256, a.n), bl(256);
  int p0 = pin[0], p1 = nin > 1 ? pin[1] : p0, p2 = nin > 2 ? pin[2] : p1;
#define QOP_C(PO_, P0_, P1_, P2_)                                                      \
  if (po == PO_ && p0 == P0_ && p1 == P1_ && p2 == P2_) {                              \
    lincomb_dispatch2<PO_, P0_, P1_, P2_>(a, nin, red, g, bl, st);                     \
    QOP_CHECK_HIP(hipGetLastError());                                                  \
    return;                                                                            \
  }
  QOP_C(P64, P64, P64, P64)
  QOP_C(P32, P32, P32, P32)
  QOP_C(P16, P16, P16, P16)
  // conversions and mixed accumulations
  QOP_C(P32, P64, P64, P64)
  QOP_C(P16, P64, P64, P64)
  QOP_C(P64, P32, P32, P32)
  QOP_C(P64, P16, P16, P16)
  QOP_C(P64, P64, P32, P32)
  QOP_C(P64, P64, P16, P16)
  QOP_C(P32, P16, P16, P16)
  QOP_C(P16, P32, P32, P32)
  QOP_C(P32, P32, P16, P16)
  QOP_C(P16, P16, P32, P32)
#undef QOP_C
  fatal(__FILE__, __LINE__, "lincomb: unsupported precision combination");
}

// ---------------------------------------------------------------------------
// dot products: partial[b][chunk] = sum conj(u) v  (2 doubles) or |u|^2
template <int PU, int PV, int MODE>  // MODE 0: |u|^2, 1: complex <u,v>
__global__ void __launch_bounds__(256)
dot_kernel(const DotArgs a) {
  const int Vh = a.Vh;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (a.status && *a.status != 0) return;
  if (a.active && !a.active[b]) return;  // inactive slot: partial rows untouched
  double sr = 0, si = 0;
  if (h < Vh) {
    typename PT<PU>::R u[6];
    vload<PU>((const char *)a.u.data + (long)b * a.ustride * sizeof(typename PT<PU>::S),
              a.u.norm ? a.u.norm + (long)b * a.nstride : nullptr, h, Vh, u);
    if constexpr (MODE == 0) {
#pragma unroll
      for (int i = 0; i < 6; i++) sr += (double)u[i] * (double)u[i];
    } else {
      typename PT<PV>::R v[6];
      vload<PV>((const char *)a.v.data + (long)b * a.vstride * sizeof(typename PT<PV>::S),
                a.v.norm ? a.v.norm + (long)b * a.nstride : nullptr, h, Vh, v);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double ur = u[2 * c], ui = u[2 * c + 1], vr = v[2 * c], vi = v[2 * c + 1];
        sr += ur * vr + ui * vi;
        si += ur * vi - ui * vr;
      }
    }
  }
  const long nchunk = (Vh + 63) / 64;
  sr = wave_sum(sr);
  if constexpr (MODE == 1) si = wave_sum(si);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nchunk) {
    if constexpr (MODE == 0) a.partial[(long)b * nchunk + (h >> 6)] = sr;
    else {
      a.partial[(long)(2 * b) * nchunk + (h >> 6)] = sr;
      a.partial[(long)(2 * b + 1) * nchunk + (h >> 6)] = si;
    }
  }
}

// Fermilab relative residual terms per site: |x|^2 == 0 ? 1 : |r|^2/|x|^2
// (lib/invert_utilities_p.c:7-26); u = r, v = x.
template <int P>
__global__ void __launch_bounds__(256)
relnorm_kernel(const DotArgs a) {
  const int Vh = a.Vh;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (a.status && *a.status != 0) return;
  double s = 0;
  if (h < Vh && (!a.active || a.active[b])) {
    typename PT<P>::R r[6], x[6];
    vload<P>((const char *)a.u.data + (long)b * a.ustride * sizeof(typename PT<P>::S),
             a.u.norm ? a.u.norm + (long)b * a.nstride : nullptr, h, Vh, r);
    vload<P>((const char *)a.v.data + (long)b * a.vstride * sizeof(typename PT<P>::S),
             a.v.norm ? a.v.norm + (long)b * a.nstride : nullptr, h, Vh, x);
    double r2 = 0, x2 = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) { r2 += (double)r[i] * r[i]; x2 += (double)x[i] * x[i]; }
    s = x2 == 0.0 ? 1.0 : r2 / x2;
  }
  const long nchunk = (Vh + 63) / 64;
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0 && (h >> 6) < nchunk) a.partial[(long)b * nchunk + (h >> 6)] = s;
}

void launch_dot(const DotArgs &a, int pu, int pv, int mode, hipStream_t st) {
  if (mode == 2) {
    dim3 g((a.Vh + 255) / 256, a.n), bl(256);
    QOP_REQUIRE(pu == pv, "relnorm: precisions must match");
    if (pu == P64) hipLaunchKernelGGL((relnorm_kernel<P64>), g, bl, 0, st, a);
    else if (pu == P32) hipLaunchKernelGGL((relnorm_kernel<P32>), g, bl, 0, st, a);
    else hipLaunchKernelGGL((relnorm_kernel<P16>), g, bl, 0, st, a);
    QOP_CHECK_HIP(hipGetLastError());
    return;
  }
  dim3 g((a.Vh + 255) / 256, a.n), bl(256);
#define QOP_D(PU_, PV_, M_) \
  if (pu == PU_ && pv == PV_ && mode == M_) { hipLaunchKernelGGL((dot_kernel<PU_, PV_, M_>), g, bl, 0, st, a); QOP_CHECK_HIP(hipGetLastError()); return; }
  QOP_D(P64, P64, 0) QOP_D(P64, P64, 1)
  QOP_D(P32, P32, 0) QOP_D(P32, P32, 1)
  QOP_D(P16, P16, 0) QOP_D(P16, P16, 1)
  if (mode == 0) {
    if (pu == P32) { hipLaunchKernelGGL((dot_kernel<P32, P32, 0>), g, bl, 0, st, a); QOP_CHECK_HIP(hipGetLastError()); return; }
  }
#undef QOP_D
  fatal(__FILE__, __LINE__, "dot: unsupported precision combination");
}

// ---------------------------------------------------------------------------
// raw (site-major [V][3][2]) <-> SoA vector, both precisions of raw
template <typename T, int P>
__global__ void raw_to_vec_kernel(const T *raw, void *data, float *norm, long Vh,
                                  long begin, long end) {
  long i = begin + blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= end) return;
  int p = i >= Vh;
  long h = i - p * Vh;
  typename PT<P>::R v[6];
#pragma unroll
  for (int c = 0; c < 6; c++) v[c] = (typename PT<P>::R)raw[6 * i + c];
  char *d = (char *)data + (size_t)p * 3 * Vh * sizeof(typename PT<P>::S);
  vstore<P>(d, norm ? norm + p * Vh : nullptr, h, Vh, v);
}

template <typename T, int P>
__global__ void vec_to_raw_kernel(T *raw, const void *data, const float *norm,
                                  long Vh, long begin, long end) {
  long i = begin + blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= end) return;
  int p = i >= Vh;
  long h = i - p * Vh;
  typename PT<P>::R v[6];
  const char *d = (const char *)data + (size_t)p * 3 * Vh * sizeof(typename PT<P>::S);
  vload<P>(d, norm ? norm + p * Vh : nullptr, h, Vh, v);
#pragma unroll
  for (int c = 0; c < 6; c++) raw[6 * i + c] = (T)v[c];
}

static void eo_range(const Lattice *lat, int eo, long &b, long &e) {
  b = eo == 1 ? lat->Vh : 0;
  e = eo == 0 ? lat->Vh : lat->V;
}

template <typename T>
static void raw_to_vec_t(const Lattice *lat, const T *raw, DevVec &v, int eo, hipStream_t st) {
  long b, e;
  eo_range(lat, eo, b, e);
  dim3 g((e - b + 255) / 256), bl(256);
  if (v.prec == P64) hipLaunchKernelGGL((raw_to_vec_kernel<T, P64>), g, bl, 0, st, raw, v.data, v.norm, lat->Vh, b, e);
  else if (v.prec == P32) hipLaunchKernelGGL((raw_to_vec_kernel<T, P32>), g, bl, 0, st, raw, v.data, v.norm, lat->Vh, b, e);
  else hipLaunchKernelGGL((raw_to_vec_kernel<T, P16>), g, bl, 0, st, raw, v.data, v.norm, lat->Vh, b, e);
  QOP_CHECK_HIP(hipGetLastError());
}
template <typename T>
static void vec_to_raw_t(const Lattice *lat, const DevVec &v, T *raw, int eo, hipStream_t st) {
  long b, e;
  eo_range(lat, eo, b, e);
  dim3 g((e - b + 255) / 256), bl(256);
  if (v.prec == P64) hipLaunchKernelGGL((vec_to_raw_kernel<T, P64>), g, bl, 0, st, raw, v.data, v.norm, lat->Vh, b, e);
  else if (v.prec == P32) hipLaunchKernelGGL((vec_to_raw_kernel<T, P32>), g, bl, 0, st, raw, v.data, v.norm, lat->Vh, b, e);
  else hipLaunchKernelGGL((vec_to_raw_kernel<T, P16>), g, bl, 0, st, raw, v.data, v.norm, lat->Vh, b, e);
  QOP_CHECK_HIP(hipGetLastError());
}

void launch_raw_to_vec(const Lattice *lat, const double *raw, DevVec &v, int eo, hipStream_t st) { raw_to_vec_t(lat, raw, v, eo, st); }
void launch_vec_to_raw(const Lattice *lat, const DevVec &v, double *raw, int eo, hipStream_t st) { vec_to_raw_t(lat, v, raw, eo, st); }
void launch_rawf_to_vec(const Lattice *lat, const float *raw, DevVec &v, int eo, hipStream_t st) { raw_to_vec_t(lat, raw, v, eo, st); }
void launch_vec_to_rawf(const Lattice *lat, const DevVec &v, float *raw, int eo, hipStream_t st) { vec_to_raw_t(lat, v, raw, eo, st); }

// ---------------------------------------------------------------------------
// plain streaming copy (16 B per lane): the calibration kernel for HBM
// traffic counters and the measured copy bandwidth of the device
__global__ void __launch_bounds__(256) copy_kernel(const float4 *__restrict__ in,
                                                   float4 *__restrict__ out, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = in[i];
}

void launch_copy(const void *in, void *out, size_t bytes, hipStream_t st) {
  long n = (long)(bytes / 16);
  hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, st, (const float4 *)in, (float4 *)out, n);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
