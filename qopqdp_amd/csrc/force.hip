// Asqtad fermion force on gfx950, the "fnmat" form of the reference
// (lib/asqtad_force_fnmat_p.c:474-546): outer products of the solution
// vectors ("mid" fields), their derivative through the fat7 + Lepage + Naik
// smearing, and force += -2 TA(U deriv^+).
//
// Each QDP field operation chain of the reference (a shift, a product, a
// shift of the product, ...) becomes one kernel that gathers what it needs
// from the neighbouring sites and writes only its own site: a staple, a
// side_force, the Naik derivative and the Lepage fix-up are one launch each,
// and no shifted copy of a field is ever stored.  Every kernel is one thread
// per site and HBM-bound (about 2 flops per byte); the neighbour reads of a
// wave are mostly L2 hits.
//
// Fields live in FP64, site-innermost: element e (row-major colour index
// 3*i+j, complex) of site s (raw order) at f[e*V + s], so each load of a
// wave is 64 consecutive 16-byte words.  Gauge and force keep the raw layout
// of the QOP API (18 doubles per site, one array per direction); the gauge
// field is copied into the site-innermost layout once per call.
#include <cstring>
#include <stdexcept>

#include "common.h"

namespace qop {
namespace {

constexpr int TPB = 256;

// Site order of the gathering kernels (x0-tiled, XCD-aware).  A slab is the
// S = Vh / L0 sites of one x0 and parity; it is cut into tiles of T sites.
// Logical blocks run tile by tile, and inside a tile over x0 and the two
// parities, so the x0 +-1 neighbours a block gathers belong to blocks about
// 2 T / TPB positions away, still in L2, and the x1..x3 neighbours mostly to
// the same or the partner block.  The hardware dispatches block b to XCD
// b % 8; XCD k runs the logical blocks [k per, (k+1) per), a contiguous part
// of that order.  -1 marks padding threads.
struct Sched {
  int S, T, nt, L0, bpt, per;  // slab, tile, tiles per slab, extent 0, blocks per tile, blocks per XCD
};
__device__ __forceinline__ int site_index(const Sched &q, int Vh) {
  const int w = (int)(blockIdx.x & 7) * q.per + (int)(blockIdx.x >> 3);
  const int sub = w % q.bpt;
  int r = w / q.bpt;
  const int par = r & 1;
  r >>= 1;
  const int x0 = r % q.L0, tile = r / q.L0;
  const int hs = tile * q.T + sub * TPB + (int)threadIdx.x;
  if (tile >= q.nt || hs >= q.S) return -1;
  return par * Vh + x0 * q.S + hs;
}

struct M3 {
  double2 e[9];
};

__device__ __forceinline__ double2 cj(double2 a) { return make_double2(a.x, -a.y); }

// op(a) op(b), op = adjoint where the flag is set
template <bool TA, bool TB>
__device__ __forceinline__ M3 mm(const M3 &a, const M3 &b) {
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 x = TA ? cj(a.e[3 * k + i]) : a.e[3 * i + k];
        const double2 y = TB ? cj(b.e[3 * j + k]) : b.e[3 * k + j];
        re = fma(x.x, y.x, fma(-x.y, y.y, re));
        im = fma(x.x, y.y, fma(x.y, y.x, im));
      }
      c.e[3 * i + j] = make_double2(re, im);
    }
  return c;
}

// c += s op(a)
template <bool TA>
__device__ __forceinline__ void axpy(M3 &c, double s, const M3 &a) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double2 x = TA ? cj(a.e[3 * j + i]) : a.e[3 * i + j];
      c.e[3 * i + j].x = fma(s, x.x, c.e[3 * i + j].x);
      c.e[3 * i + j].y = fma(s, x.y, c.e[3 * i + j].y);
    }
}

__device__ __forceinline__ M3 zero3() {
  M3 m;
#pragma unroll
  for (int e = 0; e < 9; e++) m.e[e] = make_double2(0, 0);
  return m;
}

// Element planes f + e V are uniform; the per-lane part of the address is a
// 32-bit byte offset (the volume is below 2^27 sites), so the loads and
// stores take a scalar base plus one vector offset.
__device__ __forceinline__ M3 ld(const double2 *f, long V, unsigned s) {
  M3 m;
  const unsigned off = s * (unsigned)sizeof(double2);
#pragma unroll
  for (int e = 0; e < 9; e++) m.e[e] = *(const double2 *)((const char *)(f + e * V) + off);
  return m;
}
// Non-temporal forms for a site's own matrices that a kernel reads or writes
// once (accumulators, outputs): they then take less of the L2 the
// neighbour gathers reuse.
__device__ __forceinline__ M3 ld_nt(const double2 *f, long V, unsigned s) {
  M3 m;
  const unsigned off = s * (unsigned)sizeof(double2);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    const double *p = (const double *)((const char *)(f + e * V) + off);
    m.e[e] = make_double2(__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1));
  }
  return m;
}
__device__ __forceinline__ void st_nt(double2 *f, long V, unsigned s, const M3 &m) {
  const unsigned off = s * (unsigned)sizeof(double2);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double *p = (double *)((char *)(f + e * V) + off);
    __builtin_nontemporal_store(m.e[e].x, p);
    __builtin_nontemporal_store(m.e[e].y, p + 1);
  }
}
__device__ __forceinline__ void st(double2 *f, long V, unsigned s, const M3 &m) {
  const unsigned off = s * (unsigned)sizeof(double2);
#pragma unroll
  for (int e = 0; e < 9; e++) *(double2 *)((char *)(f + e * V) + off) = m.e[e];
}

// A site's coordinates and lexicographic index (x3 fastest; 32-bit, the
// volume is below 2^31), found once per thread; neighbours then cost a few
// integer adds instead of divisions.
struct Site {
  int x[4];
  int l;
};
__device__ __forceinline__ Site site_of(const Geom &g, long s) {
  Site r;
  const int p = s >= g.Vh;
  unsigned t = 2u * (unsigned)(s - (p ? g.Vh : 0));
  const unsigned l0 = t;
  r.x[3] = (int)(t % (unsigned)g.L[3]); t /= (unsigned)g.L[3];
  r.x[2] = (int)(t % (unsigned)g.L[2]); t /= (unsigned)g.L[2];
  r.x[1] = (int)(t % (unsigned)g.L[1]); t /= (unsigned)g.L[1];
  r.x[0] = (int)t;
  r.l = (int)l0;
  if (((r.x[0] + r.x[1] + r.x[2] + r.x[3]) & 1) != p) {
    r.x[3] += 1;  // L3 is even: x3 + 1 stays inside
    r.l += 1;
  }
  return r;
}

// raw index of x + a mu + b nu (mu != nu or b == 0, |a|, |b| <= 3), periodic.
// The extents are even (at least 2), so two conditional wraps suffice and
// the parity of the neighbour is that of x + a + b.
__device__ __forceinline__ int wrap(int y, int L) {
  y = y >= L ? y - L : y;
  y = y >= L ? y - L : y;
  y = y < 0 ? y + L : y;
  y = y < 0 ? y + L : y;
  return y;
}
__device__ __forceinline__ unsigned nbr(const Geom &g, const Site &x, int mu, int a, int nu, int b) {
  int l = x.l + (wrap(x.x[mu] + a, g.L[mu]) - x.x[mu]) * g.stride[mu];
  if (b) l += (wrap(x.x[nu] + b, g.L[nu]) - x.x[nu]) * g.stride[nu];
  const int p = (x.x[0] + x.x[1] + x.x[2] + x.x[3] + a + b) & 1;
  return (unsigned)(l >> 1) + (p ? (unsigned)g.Vh : 0u);
}

// A field given as ca A + B (B may be null): the bottom input of side_force,
// which the reference forms in a temporary (lib/asqtad_force_fnmat_p.c:290-300).
struct Lin {
  const double2 *a;
  double ca;
  const double2 *b;  // used when the kernel is instantiated with HASB
};
template <bool HASB>
__device__ __forceinline__ M3 ld(const Lin &f, long V, unsigned s) {
  M3 m = ld(f.a, V, s);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    m.e[e].x *= f.ca;
    m.e[e].y *= f.ca;
  }
  if constexpr (HASB) {
    const M3 t = ld(f.b, V, s);
#pragma unroll
    for (int e = 0; e < 9; e++) {
      m.e[e].x += t.e[e].x;
      m.e[e].y += t.e[e].y;
    }
  }
  return m;
}

// Four field pointers passed by value (one per direction).
struct Ptr4 {
  const double2 *p[4];
};

// raw [V][18] -> site-innermost [9][V].  A block's TPB sites are one
// contiguous run of 9 TPB complex numbers in the raw layout; it is read
// whole into LDS and written out element-plane by element-plane, so both
// sides are contiguous.
__global__ __launch_bounds__(TPB) void to_soa_kernel(const double2 *raw, double2 *f, long V) {
  __shared__ double2 t[9 * TPB];
  const long s0 = (long)blockIdx.x * TPB;
  const int n = V - s0 < TPB ? (int)(V - s0) : TPB;
  const double2 *src = raw + 9 * s0;
  for (int i = threadIdx.x; i < 9 * n; i += TPB) t[i] = src[i];
  __syncthreads();
  if ((int)threadIdx.x < n) {
#pragma unroll
    for (int e = 0; e < 9; e++) f[e * V + s0 + threadIdx.x] = t[9 * threadIdx.x + e];
  }
}

// o = co o + ca A + cb B; co == 0 does not read o, null A/B are zero
__global__ __launch_bounds__(TPB) void lincomb_kernel(double2 *o, double co, double ca, const double2 *A, double cb,
                               const double2 *B, long n) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double2 r = make_double2(0, 0);
  if (co != 0) {
    r = o[k];
    r.x *= co;
    r.y *= co;
  }
  if (A) {
    r.x = fma(ca, A[k].x, r.x);
    r.y = fma(ca, A[k].y, r.y);
  }
  if (B) {
    r.x = fma(cb, B[k].x, r.x);
    r.y = fma(cb, B[k].y, r.y);
  }
  o[k] = r;
}

// QOP_get_mid (lib/asqtad_force_fnmat_p.c:4-44) with its factor (scale: -1
// in the asqtad force, :491-503; +1 in the HISQ force,
// lib/hisq_force_fnmat_p.c:63-76) and the negation on odd sites:
// mid_s(x) = scale (-1)^p(x) sum_t eps_t x_t(x) x_t(x + dist_s)^+,
// dist_s = s-th direction (s < 4) or three times it (s >= 4).  Vectors are
// full-lattice [parity][colour][site] complex.
template <typename T2>
__global__ __launch_bounds__(TPB) void mid_kernel(Geom g, Sched q, long V, const T2 *const *xs,
                                                  const double *eps, int n, int nmid, double scale,
                                                  double2 *mid) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const long Vh = g.Vh;
  const double sg = s >= Vh ? -scale : scale;
  auto vidx = [&](long i, int c) { return i + (i >= Vh ? 2 * Vh : 0) + c * Vh; };
  // four mids per pass (one- and three-hop of two directions when nmid = 8),
  // so x_t(x) is loaded once per pass instead of once per mid
  for (int g0 = 0; g0 < nmid; g0 += 4) {
    int mi[4];
    unsigned j[4];
    M3 acc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      mi[k] = nmid == 8 ? (g0 / 2 + (k >> 1)) + 4 * (k & 1) : g0 + k;
      j[k] = nbr(g, x, mi[k] & 3, mi[k] < 4 ? 1 : 3, 0, 0);
      acc[k] = zero3();
    }
    for (int t = 0; t < n; t++) {
      const T2 *v = xs[t];
      double2 a[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const T2 va = v[vidx(s, c)];
        a[c] = make_double2(va.x, va.y);
      }
      const double w = sg * eps[t];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        double2 b[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const T2 vb = v[vidx(j[k], c)];
          b[c] = make_double2(vb.x, vb.y);
        }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const double re = fma(a[r].x, b[c].x, a[r].y * b[c].y);
            const double im = fma(a[r].y, b[c].x, -a[r].x * b[c].y);
            acc[k].e[3 * r + c].x = fma(w, re, acc[k].e[3 * r + c].x);
            acc[k].e[3 * r + c].y = fma(w, im, acc[k].e[3 * r + c].y);
          }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) st_nt(mid + (long)mi[k] * 9 * V, V, s, acc[k]);
  }
}

// staple (lib/asqtad_force_fnmat_p.c:93-127) with its scaling folded in:
// out(x) = [out(x) if acc] + sc [K(x) I(x+nu) K(x+mu)^+ + K(x-nu)^+ I(x-nu) K(x-nu+mu)]
//          + te E(x) [+ the same staple of I2, K2 along nu2 if TWO]
// I is the field along mu, K the link along nu.  TWO computes the two
// 7-staples of a (sig, rho) pair in one pass, in the order in which the
// reference accumulates them (:252-262).
__device__ __forceinline__ void add_staple(M3 &r, double sc, const double2 *I, const double2 *K, long V,
                                           const Geom &g, const Site &x, unsigned s, int mu, int nu) {
  const unsigned xpnu = nbr(g, x, nu, 1, mu, 0), xpmu = nbr(g, x, mu, 1, nu, 0);
  const unsigned xmnu = nbr(g, x, nu, -1, mu, 0), xmnupmu = nbr(g, x, nu, -1, mu, 1);
  {
    const M3 t = mm<false, true>(ld(I, V, xpnu), ld(K, V, xpmu));
    axpy<false>(r, sc, mm<false, false>(ld(K, V, s), t));
  }
  {
    const M3 t = mm<true, false>(ld(K, V, xmnu), ld(I, V, xmnu));
    axpy<false>(r, sc, mm<false, false>(t, ld(K, V, xmnupmu)));
  }
}
template <bool ACC, bool HASE, bool TWO = false>
__global__ __launch_bounds__(TPB) void staple_kernel(Geom g, Sched q, long V, double2 *out, double sc, const double2 *I,
                                                     const double2 *K, int mu, int nu, double te, const double2 *E,
                                                     const double2 *I2, const double2 *K2, int nu2) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  M3 r = ACC ? ld_nt(out, V, s) : zero3();
  add_staple(r, sc, I, K, V, g, x, s, mu, nu);
  if constexpr (HASE) axpy<false>(r, te, ld(E, V, s));
  if constexpr (TWO) {
    __builtin_amdgcn_sched_barrier(0);
    add_staple(r, sc, I2, K2, V, g, x, s, mu, nu2);
  }
  st_nt(out, V, s, r);
}

// side_force (lib/asqtad_force_fnmat_p.c:129-199) at one site:
// force(x) += T(x) S(x+mu) B(x+nu)^+ + B(x) S(x+mu) T(x+nu)^+
//           + B(x-mu)^+ S(x-mu) T(x-mu+nu) + T(x-mu)^+ S(x-mu) B(x-mu+nu)
// stpl(x) = [stpl(x) if acc] + S(x) B(x+nu) S(x+mu)^+ + S(x-nu)^+ B(x-nu) S(x-nu+mu)
// with B = bot, S = side (the link along nu), T = top.
// EPI: what is done with the staple S at the end.
// EPI_STORE: stpl = S.
// EPI_MIX (the 7-staple/Lepage side_force): S is mixed as the reference does
// next (lib/asqtad_force_fnmat_p.c:268-277): mid5 = [mid5 if m5acc] + cl S
// (mid5 may be null), stpl = c7 S + c5 ms.
// EPI_FIN (the last 3-staple side_force of a direction sig): S is not
// stored but added to the derivative of sig (:309-314):
// d = d + S [+ c1 ms if ms].
enum { EPI_STORE = 0, EPI_MIX = 1, EPI_FIN = 2 };
struct Mix {
  double2 *mid5;  // EPI_FIN: the derivative d
  double cl, c7, c5;  // EPI_FIN: c5 = c1
  const double2 *ms;
  int m5acc;
};
template <bool ACC, bool HASB, int EPI = EPI_STORE>
__global__ __launch_bounds__(TPB) void side_force_kernel(Geom g, Sched q, long V, double2 *force, double2 *stpl,
                                                         Lin bot, const double2 *side,
                                                         const double2 *top, int mu, int nu, Mix mx) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const unsigned xpmu = nbr(g, x, mu, 1, nu, 0), xpnu = nbr(g, x, nu, 1, mu, 0);
  // SB() keeps the scheduler from hoisting every matrix load to the top:
  // with all 14 in flight a thread needs ~500 registers (one wave per SIMD).
  // B(x+nu) S(x+mu)^+ and S(x+mu) are loaded once and serve both the force
  // terms at +mu and the staple; the staple is formed last, after the force
  // is stored, so its epilogue (EPI) has the registers of the force terms.
#define SB() __builtin_amdgcn_sched_barrier(0)
  M3 f, bs;
  {
    const M3 Smu = ld(side, V, xpmu);
    bs = mm<false, true>(ld<HASB>(bot, V, xpnu), Smu);  // B(x+nu) S(x+mu)^+
    SB();
    f = mm<false, true>(ld(top, V, s), bs);  // T S(x+mu) B(x+nu)^+
    SB();
    const M3 b2 = mm<false, false>(ld<HASB>(bot, V, s), Smu);  // B(x) S(x+mu)
    SB();
    axpy<false>(f, 1.0, mm<false, true>(b2, ld(top, V, xpnu)));
  }
  SB();
  {
    const unsigned xmmu = nbr(g, x, mu, -1, nu, 0), xmmupnu = nbr(g, x, mu, -1, nu, 1);
    const M3 Sm = ld(side, V, xmmu);
    const M3 t1 = mm<true, false>(ld<HASB>(bot, V, xmmu), Sm);  // B(x-mu)^+ S(x-mu)
    SB();
    axpy<false>(f, 1.0, mm<false, false>(t1, ld(top, V, xmmupnu)));
    SB();
    const M3 t2 = mm<true, false>(ld(top, V, xmmu), Sm);  // T(x-mu)^+ S(x-mu)
    SB();
    axpy<false>(f, 1.0, mm<false, false>(t2, ld<HASB>(bot, V, xmmupnu)));
  }
  SB();
  axpy<false>(f, 1.0, ld_nt(force, V, s));
  st_nt(force, V, s, f);
  SB();
  {
    const unsigned xmnu = nbr(g, x, nu, -1, mu, 0), xmnupmu = nbr(g, x, nu, -1, mu, 1);
    M3 sp = mm<false, false>(ld(side, V, s), bs);
    SB();
    const M3 t = mm<true, false>(ld(side, V, xmnu), ld<HASB>(bot, V, xmnu));
    SB();
    axpy<false>(sp, 1.0, mm<false, false>(t, ld(side, V, xmnupmu)));
    if constexpr (ACC) axpy<false>(sp, 1.0, ld_nt(stpl, V, s));
    if constexpr (EPI == EPI_FIN) {
      SB();
      M3 d = ld_nt(mx.mid5, V, s);
      axpy<false>(d, 1.0, sp);
      if (mx.ms) {
        SB();
        axpy<false>(d, mx.c5, ld(mx.ms, V, s));
      }
      st_nt(mx.mid5, V, s, d);
      return;
    }
    if constexpr (EPI == EPI_MIX) {
      SB();
      if (mx.mid5) {
        M3 a = mx.m5acc ? ld_nt(mx.mid5, V, s) : zero3();
#pragma unroll
        for (int e = 0; e < 9; e++) {
          a.e[e].x = fma(mx.cl, sp.e[e].x, a.e[e].x);
          a.e[e].y = fma(mx.cl, sp.e[e].y, a.e[e].y);
        }
        st_nt(mx.mid5, V, s, a);
      }
      SB();
      const M3 m = ld(mx.ms, V, s);
#pragma unroll
      for (int e = 0; e < 9; e++) {
        sp.e[e].x = fma(mx.c7, sp.e[e].x, mx.c5 * m.e[e].x);
        sp.e[e].y = fma(mx.c7, sp.e[e].y, mx.c5 * m.e[e].y);
      }
    }
    st_nt(stpl, V, s, sp);
  }
#undef SB
}

// Naik part of QOP_asqtad_deriv (lib/asqtad_force_fnmat_p.c:395-442):
// D(x) += cn [m(x) U(x+2mu)^+ U(x+mu)^+ + U(x-mu)^+ m(x-mu) U(x+mu)^+
//             + U(x-mu)^+ U(x-2mu)^+ m(x-2mu)]
__global__ __launch_bounds__(TPB) void naik_deriv_kernel(Geom g, Sched q, long V, double2 *D, const double2 *U, const double2 *m,
                                  double cn, int mu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const unsigned p1 = nbr(g, x, mu, 1, 0, 0), p2 = nbr(g, x, mu, 2, 0, 0);
  const unsigned m1 = nbr(g, x, mu, -1, 0, 0), m2 = nbr(g, x, mu, -2, 0, 0);
  const M3 Up1 = ld(U, V, p1), Um1 = ld(U, V, m1);
  M3 f = mm<false, true>(ld(m, V, s), mm<false, false>(Up1, ld(U, V, p2)));  // m [U(x+mu) U(x+2mu)]^+
  const M3 a = mm<true, false>(Um1, ld(m, V, m1));
  axpy<false>(f, 1.0, mm<false, true>(a, Up1));
  const M3 b = mm<false, false>(ld(U, V, m2), Um1);  // [U(x-2mu) U(x-mu)]^+ = U(x-mu)^+ U(x-2mu)^+
  axpy<false>(f, 1.0, mm<true, false>(b, ld(m, V, m2)));
  M3 d = ld_nt(D, V, s);
  axpy<false>(d, cn, f);
  st_nt(D, V, s, d);
}

// Lepage fix-up (lib/asqtad_force_fnmat_p.c:317-339), first half:
// acc(x) = sum_{nu != mu} [q_nu(x) + q_nu(x)^+], q_nu(x) = m_nu(x) U_nu(x)^+ - m_nu(x-nu)^+ U_nu(x-nu)
__global__ __launch_bounds__(TPB) void lepage_acc_kernel(Geom g, Sched q, long V, double2 *acc, Ptr4 U,
                                                         Ptr4 m, int mu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  M3 r = zero3();
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const unsigned b = nbr(g, x, nu, -1, 0, 0);
    M3 t = mm<false, true>(ld(m.p[nu], V, s), ld(U.p[nu], V, s));
    axpy<false>(t, -1.0, mm<true, false>(ld(m.p[nu], V, b), ld(U.p[nu], V, b)));
    axpy<false>(r, 1.0, t);
    axpy<true>(r, 1.0, t);
  }
  st(acc, V, s, r);
}

// second half: D(x) += cf [acc(x) U(x) - U(x) acc(x+mu)]
__global__ __launch_bounds__(TPB) void lepage_apply_kernel(Geom g, Sched q, long V, double2 *D, const double2 *U,
                                    const double2 *acc, double cf, int mu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const M3 u = ld(U, V, s);
  M3 t = mm<false, false>(ld(acc, V, s), u);
  axpy<false>(t, -1.0, mm<false, false>(u, ld(acc, V, nbr(g, x, mu, 1, 0, 0))));
  M3 d = ld_nt(D, V, s);
  axpy<false>(d, cf, t);
  st_nt(D, V, s, d);
}

// force (raw layout) += fs TA(U D^+), TA the traceless antihermitian part:
// fs = -2 in the asqtad force (lib/asqtad_force_fnmat_p.c:531-539), +2 in the
// HISQ force (factor 2 of lib/hisq_force_fnmat_p.c:171-181, then :211-219)
__global__ __launch_bounds__(TPB) void force_kernel(long V, double2 *F, const double2 *U, const double2 *D,
                                                    double fs) {
  // the raw force of the block's TPB sites is one contiguous run, moved
  // through LDS so that it is read and written contiguously
  __shared__ double2 ft[9 * TPB];
  const long s0 = (long)blockIdx.x * TPB;
  const int n = V - s0 < TPB ? (int)(V - s0) : TPB;
  double2 *Fb = F + 9 * s0;
  for (int i = threadIdx.x; i < 9 * n; i += TPB) ft[i] = Fb[i];
  __syncthreads();
  if ((int)threadIdx.x < n) {
    const long s = s0 + threadIdx.x;
    const M3 a = mm<false, true>(ld(U, V, s), ld(D, V, s));
    M3 t;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const double2 p = a.e[3 * i + j], q = a.e[3 * j + i];
        t.e[3 * i + j] = make_double2(0.5 * (p.x - q.x), 0.5 * (p.y + q.y));
      }
    const double tr = (t.e[0].y + t.e[4].y + t.e[8].y) / 3.0;
    t.e[0].y -= tr;
    t.e[4].y -= tr;
    t.e[8].y -= tr;
#pragma unroll
    for (int e = 0; e < 9; e++) {
      double2 f = ft[9 * threadIdx.x + e];
      f.x = fma(fs, t.e[e].x, f.x);
      f.y = fma(fs, t.e[e].y, f.y);
      ft[9 * threadIdx.x + e] = f;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 9 * n; i += TPB) Fb[i] = ft[i];
}

// ---------------------------------------------------------------------------
// Asqtad fattening (lib/asqtad_dslash_p.c:391-633, compute_gen_staple /
// make_imp_links) on the same site-innermost fields and site order: each
// generalised staple S(x) = K(x) I(x+nu) K(x+mu)^+ + K(x-nu)^+ I(x-nu) K(x-nu+mu)
// (I the field along mu, K the link along nu) is one kernel that adds c S
// to the fat link and, when the next smearing level needs it, stores S.
template <bool STORE>
__global__ __launch_bounds__(TPB) void fat_staple_kernel(Geom g, Sched q, long V, double2 *out, double2 *fat,
                                                         double c, const double2 *I, const double2 *K,
                                                         int mu, int nu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const unsigned xpnu = nbr(g, x, nu, 1, mu, 0), xpmu = nbr(g, x, mu, 1, nu, 0);
  const unsigned xmnu = nbr(g, x, nu, -1, mu, 0), xmnupmu = nbr(g, x, nu, -1, mu, 1);
  M3 r = mm<false, false>(ld(K, V, s), mm<false, true>(ld(I, V, xpnu), ld(K, V, xpmu)));
  {
    const M3 t = mm<true, false>(ld(K, V, xmnu), ld(I, V, xmnu));
    axpy<false>(r, 1.0, mm<false, false>(t, ld(K, V, xmnupmu)));
  }
  if constexpr (STORE) st_nt(out, V, s, r);
  M3 f = ld_nt(fat, V, s);
  axpy<false>(f, c, r);
  st_nt(fat, V, s, f);
}

// Naik link L(x) = c U(x) U(x+mu) U(x+2mu) (lib/asqtad_dslash_p.c:600-620)
__global__ __launch_bounds__(TPB) void naik_link_kernel(Geom g, Sched q, long V, double2 *L, const double2 *U,
                                                        double c, int mu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const M3 t = mm<false, false>(ld(U, V, nbr(g, x, mu, 1, 0, 0)), ld(U, V, nbr(g, x, mu, 2, 0, 0)));
  M3 o = mm<false, false>(ld(U, V, s), t);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    o.e[e].x *= c;
    o.e[e].y *= c;
  }
  st(L, V, s, o);
}

// site-innermost [9][V] -> raw [V][18], through LDS as to_soa_kernel
__global__ __launch_bounds__(TPB) void from_soa_kernel(const double2 *f, double2 *raw, long V) {
  __shared__ double2 t[9 * TPB];
  const long s0 = (long)blockIdx.x * TPB;
  const int n = V - s0 < TPB ? (int)(V - s0) : TPB;
  if ((int)threadIdx.x < n) {
#pragma unroll
    for (int e = 0; e < 9; e++) t[9 * threadIdx.x + e] = f[e * V + s0 + threadIdx.x];
  }
  __syncthreads();
  double2 *dst = raw + 9 * s0;
  for (int i = threadIdx.x; i < 9 * n; i += TPB) dst[i] = t[i];
}

// Derivative of the U(3) projection W = V (V^+V)^-1/2 contracted with F2,
// the derivative accumulated with respect to W (lib/hisq_reunit1_p.c:366-433
// and the derivative of lib/hisq_reunit_utilities_p.c:301-640, computed here
// in the eigenbasis of Q = V^+V instead of the reference's closed form):
// with Q = E diag(l) E^+ from a one-sided Jacobi rotation of V (a = V E has
// orthogonal columns, l_j = |a_j|^2), s = sqrt(l),
//   Y = Q^-1/2 F2^+ + (N + N^+) V^+,
//   N = E [G o (E^+ F2^+ V E)] E^+,  G_ij = -1 / (s_i s_j (s_i + s_j)),
// i.e. Y(n,m) = d[tr(W F2^+) + tr(W^+ F2)]/dV_mn; G has no division by
// eigenvalue differences, so equal eigenvalues need no special case.  The
// force filter (hisq_reunit_utilities_p.c:437-456) adds f to every eigenvalue
// when the smallest is below f.  Writes the level-1 mid Y^+ (site-innermost)
// and counts the filtered links.
template <int P, int Q>
__device__ __forceinline__ void jacobi_rot(M3 &a, M3 &E, double &worst) {
  double al = 0, be = 0, gr = 0, gi = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double2 x = a.e[3 * k + P], y = a.e[3 * k + Q];
    al = fma(x.x, x.x, fma(x.y, x.y, al));
    be = fma(y.x, y.x, fma(y.y, y.y, be));
    gr = fma(x.x, y.x, fma(x.y, y.y, gr));
    gi = fma(x.x, y.y, fma(-x.y, y.x, gi));
  }
  const double g = sqrt(gr * gr + gi * gi);
  const double ab = sqrt(al * be);
  if (g == 0 || g <= 1e-17 * ab) return;
  worst = fmax(worst, g / ab);
  const double cph = gr / g, sph = gi / g;
  const double zeta = (be - al) / (2 * g);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
  const double c = 1 / sqrt(1 + t * t), sn = c * t;
  // columns P, Q:  p' = c p - s e^{-i phi} q,  q' = s e^{i phi} p + c q
#pragma unroll
  for (int m = 0; m < 2; m++) {
    M3 &X = m ? E : a;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double2 pv = X.e[3 * k + P], qv = X.e[3 * k + Q];
      X.e[3 * k + P] = make_double2(c * pv.x - sn * (cph * qv.x + sph * qv.y),
                                    c * pv.y - sn * (cph * qv.y - sph * qv.x));
      X.e[3 * k + Q] = make_double2(sn * (cph * pv.x - sph * pv.y) + c * qv.x,
                                    sn * (cph * pv.y + sph * pv.x) + c * qv.y);
    }
  }
}

__global__ __launch_bounds__(TPB) void u3_deriv_kernel(long V, const double2 *Vraw, const double2 *F2,
                                                       double2 *mid, double filter, int *nfilt) {
  const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (s >= V) return;
  M3 v;
#pragma unroll
  for (int e = 0; e < 9; e++) v.e[e] = Vraw[9 * s + e];
  M3 a = v, E = zero3();
  E.e[0].x = E.e[4].x = E.e[8].x = 1.0;
  for (int sweep = 0; sweep < 40; sweep++) {
    double worst = 0;
    jacobi_rot<0, 1>(a, E, worst);
    jacobi_rot<0, 2>(a, E, worst);
    jacobi_rot<1, 2>(a, E, worst);
    if (worst < 1e-16) break;
  }
  double l[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    l[j] = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) l[j] = fma(a.e[3 * k + j].x, a.e[3 * k + j].x, fma(a.e[3 * k + j].y, a.e[3 * k + j].y, l[j]));
  }
  if (filter > 0 && (l[0] < filter || l[1] < filter || l[2] < filter)) {
    l[0] += filter;
    l[1] += filter;
    l[2] += filter;
    atomicAdd(nfilt, 1);
  }
  double sq[3];
#pragma unroll
  for (int j = 0; j < 3; j++) sq[j] = sqrt(l[j]);
  const M3 F = ld(F2, V, (unsigned)s);
  // Q^-1/2 = E diag(1/s) E^+
  M3 t = E;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      t.e[3 * i + j].x /= sq[j];
      t.e[3 * i + j].y /= sq[j];
    }
  const M3 Qm = mm<false, true>(t, E);
  M3 Mp = mm<false, false>(mm<true, false>(E, mm<true, false>(F, v)), E);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double gm = -1.0 / (sq[i] * sq[j] * (sq[i] + sq[j]));
      Mp.e[3 * i + j].x *= gm;
      Mp.e[3 * i + j].y *= gm;
    }
  const M3 N = mm<false, true>(mm<false, false>(E, Mp), E);
  M3 D = N;
  axpy<true>(D, 1.0, N);
  M3 Y = mm<false, true>(Qm, F);
  axpy<false>(Y, 1.0, mm<false, true>(D, v));
  // mid = Y^+
  M3 o;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) o.e[3 * i + j] = cj(Y.e[3 * j + i]);
  st(mid, V, (unsigned)s, o);
}

inline dim3 grid(long n) { return dim3((unsigned)((n + TPB - 1) / TPB)); }
// The schedule for a lattice: tiles of up to 4096 sites (whole blocks).
inline Sched make_sched(const Lattice *lat) {
  Sched q;
  q.L0 = lat->L[0];
  q.S = (int)(lat->Vh / q.L0);
  const int Sb = (q.S + TPB - 1) / TPB * TPB;
  q.T = Sb < 4096 ? Sb : 4096;
  q.nt = (q.S + q.T - 1) / q.T;
  q.bpt = q.T / TPB;
  const long W = (long)q.nt * q.L0 * 2 * q.bpt;
  q.per = (int)((W + 7) / 8);
  return q;
}

}  // namespace

// Reference flop tallies: lib/asqtad_force_fnmat_p.c:117, 185, 41, 338, 445, 544.
constexpr double STAPLE_FLOPS = 3 * 198 + 216 + 18;
constexpr double SIDE_FORCE_FLOPS = 7 * 198 + 3 * 216 + 3 * 18;

namespace {

// One force evaluation: launch geometry, site-innermost workspace and the
// device table of vector pointers and eps.  nfields: fields besides the
// derivative scratch (stpl3[4], mid5[4], stpl5, mid3).
class ForceRun {
 public:
  ForceRun(const Lattice *lat, int nfields, const double *eps, const DevVec *const *xs, int n,
           hipStream_t st)
      : g_(lat->g), V_(lat->V), n9_(9 * lat->V), st_(st), sq_(make_sched(lat)),
        b_(TPB), gs_(grid(lat->V)), ge_(grid(9 * lat->V)), gx_(8 * sq_.per) {
    if (V_ >= (1L << 27))
      throw std::runtime_error("fermion force: lattice volume of 2^27 sites or more (32-bit offsets)");
    prec_ = n > 0 ? xs[0]->prec : P64;
    for (int t = 0; t < n; t++)
      if (xs[t]->prec != prec_ || xs[t]->lat != lat || (prec_ != P64 && prec_ != P32))
        throw std::runtime_error("fermion force: vectors must share the lattice and be FP64 or FP32");
    dev_alloc(&buf_.p, (size_t)(nfields + 10) * n9_ * sizeof(double2));
    next_ = buf_.p;
    for (auto &q : stpl3_) q = take();
    for (auto &q : mid5_) q = take();
    stpl5_ = take();
    mid3_ = take();
    const size_t ntab = (size_t)2 * n + 1;
    dev_alloc(&tab_.p, ntab * sizeof(double));
    std::vector<double> h(ntab);
    for (int t = 0; t < n; t++) {
      const void *d = xs[t]->data;
      memcpy(&h[t], &d, sizeof d);
      h[n + t] = eps[t];
    }
    QOP_CHECK_HIP(hipMemcpy(tab_.p, h.data(), ntab * sizeof(double), hipMemcpyHostToDevice));
    n_ = n;
  }
  double2 *take() {
    double2 *r = next_;
    next_ += n9_;
    return r;
  }
  double flops_per_site = 0;

  void to_soa(const double *raw, double2 *f) {
    hipLaunchKernelGGL(to_soa_kernel, gs_, b_, 0, st_, (const double2 *)raw, f, V_);
  }
  // o = co o + ca A + cb B
  void lincomb(double2 *o, double co, double ca, const double2 *A, double cb, const double2 *B) {
    hipLaunchKernelGGL(lincomb_kernel, ge_, b_, 0, st_, o, co, ca, A, cb, B, n9_);
  }
  // mid fields of terms [t0, t0 + cnt): nmid = 4 (one-hop) or 8 (and three-hop)
  void mids(int t0, int cnt, int nmid, double scale, double2 *mid) {
    const double *tab = tab_.p;
    if (prec_ == P64)
      hipLaunchKernelGGL(mid_kernel<double2>, gx_, b_, 0, st_, g_, sq_, V_,
                         (const double2 *const *)(tab + t0), tab + n_ + t0, cnt, nmid, scale, mid);
    else
      hipLaunchKernelGGL(mid_kernel<float2>, gx_, b_, 0, st_, g_, sq_, V_,
                         (const float2 *const *)(tab + t0), tab + n_ + t0, cnt, nmid, scale, mid);
    flops_per_site += (54 + 36) * 4.0 * cnt * (nmid / 4);
  }
  // QOP_asqtad_deriv (lib/asqtad_force_fnmat_p.c:201-464): D += the
  // derivative of the smearing of the links Ul with coefficients coef for
  // the mids (midN: three-hop mids, used when coef[5] != 0).
  void deriv(double2 *const Ul[4], double2 *D[4], const double coef[6], double2 *const mid[4],
             double2 *const midN[4]);
  // force (raw) += fs TA(Ul D^+)
  void final(double *F_raw[4], double2 *const Ul[4], double2 *const D[4], double fs) {
    for (int mu = 0; mu < 4; mu++)
      hipLaunchKernelGGL(force_kernel, gs_, b_, 0, st_, V_, (double2 *)F_raw[mu], (const double2 *)Ul[mu],
                         (const double2 *)D[mu], fs);
    flops_per_site += 4 * (198 + 24 + 18);
  }
  void u3_deriv(const double *Vraw, const double2 *F2, double2 *mid, double filter, int *nfilt) {
    hipLaunchKernelGGL(u3_deriv_kernel, gs_, b_, 0, st_, V_, (const double2 *)Vraw, F2, mid, filter, nfilt);
  }
  void finish() {
    QOP_CHECK_HIP(hipGetLastError());
    QOP_CHECK_HIP(hipStreamSynchronize(st_));
  }
  long V() const { return V_; }

 private:
  void staple(double2 *out, int acc, double sc, const double2 *I, const double2 *K, int mu, int nu,
              double te, const double2 *E) {
    const double2 *n2 = nullptr;
    if (acc && E)
      hipLaunchKernelGGL((staple_kernel<true, true>), gx_, b_, 0, st_, g_, sq_, V_, out, sc, I, K, mu, nu, te, E, n2, n2, 0);
    else if (acc)
      hipLaunchKernelGGL((staple_kernel<true, false>), gx_, b_, 0, st_, g_, sq_, V_, out, sc, I, K, mu, nu, te, E, n2, n2, 0);
    else if (E)
      hipLaunchKernelGGL((staple_kernel<false, true>), gx_, b_, 0, st_, g_, sq_, V_, out, sc, I, K, mu, nu, te, E, n2, n2, 0);
    else
      hipLaunchKernelGGL((staple_kernel<false, false>), gx_, b_, 0, st_, g_, sq_, V_, out, sc, I, K, mu, nu, te, E, n2, n2, 0);
  }
  // out = sc S(I, K, nu) [+ te E] + sc S(I2, K2, nu2)
  void staple2(double2 *out, double sc, const double2 *I, const double2 *K, int mu, int nu, double te,
               const double2 *E, const double2 *I2, const double2 *K2, int nu2) {
    if (E)
      hipLaunchKernelGGL((staple_kernel<false, true, true>), gx_, b_, 0, st_, g_, sq_, V_, out, sc, I, K, mu, nu, te, E, I2,
                         K2, nu2);
    else
      hipLaunchKernelGGL((staple_kernel<false, false, true>), gx_, b_, 0, st_, g_, sq_, V_, out, sc, I, K, mu, nu, te, E, I2,
                         K2, nu2);
  }
  void side_force(double2 *force, Lin bot, const double2 *side, const double2 *top, int mu, int nu,
                  double2 *stpl, int acc) {
    const Mix mx{};
    if (acc && bot.b)
      hipLaunchKernelGGL((side_force_kernel<true, true>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top, mu, nu, mx);
    else if (acc)
      hipLaunchKernelGGL((side_force_kernel<true, false>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top, mu, nu, mx);
    else if (bot.b)
      hipLaunchKernelGGL((side_force_kernel<false, true>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top, mu, nu, mx);
    else
      hipLaunchKernelGGL((side_force_kernel<false, false>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top, mu, nu, mx);
  }
  // side_force with a single bottom field and the mix folded in (stpl = c7 S + c5 ms)
  void side_force_mix(double2 *force, Lin bot, const double2 *side, const double2 *top, int mu, int nu,
                      double2 *stpl, const Mix &mx) {
    hipLaunchKernelGGL((side_force_kernel<false, false, EPI_MIX>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top,
                       mu, nu, mx);
  }
  // accumulating side_force whose staple goes into d (+ c1 ms) instead of a field
  void side_force_fin(double2 *force, Lin bot, const double2 *side, const double2 *top, int mu, int nu,
                      double2 *stpl, double2 *d, double c1, const double2 *ms) {
    const Mix mx{d, 0.0, 0.0, c1, ms, 0};
    if (bot.b)
      hipLaunchKernelGGL((side_force_kernel<true, true, EPI_FIN>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top,
                         mu, nu, mx);
    else
      hipLaunchKernelGGL((side_force_kernel<true, false, EPI_FIN>), gx_, b_, 0, st_, g_, sq_, V_, force, stpl, bot, side, top,
                         mu, nu, mx);
  }

  Geom g_;
  long V_, n9_;
  hipStream_t st_;
  Sched sq_;
  dim3 b_, gs_, ge_, gx_;
  int prec_ = P64, n_ = 0;
  DevBuf<double2> buf_;
  DevBuf<double> tab_;
  double2 *next_ = nullptr;
  double2 *stpl3_[4] = {}, *mid5_[4] = {}, *stpl5_ = nullptr, *mid3_ = nullptr;
};

void ForceRun::deriv(double2 *const U[4], double2 *D[4], const double coef[6], double2 *const mid[4],
                     double2 *const midN[4]) {
  double coef1 = coef[0], coef3 = coef[1], coef5 = coef[2], coef7 = coef[3], coefL = coef[4];
  coef1 -= 6 * coefL;
  const bool have5 = coef5 != 0 || coef7 != 0 || coefL != 0;
  const bool have3 = coef3 != 0 || have5;
  double2 **stpl3 = stpl3_, **mid5 = mid5_, *stpl5 = stpl5_, *mid3 = mid3_;
  double &nfl = flops_per_site;
  // QOP_fat_deriv (lib/asqtad_force_fnmat_p.c:236-316)
  for (int sig = 0; sig < 4; sig++) {
    if (have5) {
      for (int mu = 0; mu < 4; mu++) {
        if (mu == sig) continue;
        staple(stpl3[mu], 0, 1.0, U[sig], U[mu], sig, mu, 0.0, nullptr);
        nfl += STAPLE_FLOPS;
      }
      // mid5[nu] starts at zero (:243-247): its first update overwrites
      bool m5set[4] = {false, false, false, false};
      for (int rho = 0; rho < 4; rho++) {
        if (rho == sig) continue;
        // stpl5 = coef7 sum 7-staples + coefL stpl3[rho]
        if (coef7 != 0) {
          // the two (mu, nu) pairs of directions other than sig and rho
          int pm[2], pn[2], np = 0;
          for (int mu = 0; mu < 4; mu++) {
            if (mu == sig || mu == rho) continue;
            for (int nu = 0; nu < 4; nu++) {
              if (nu == sig || nu == rho || nu == mu) continue;
              pm[np] = mu;
              pn[np++] = nu;
              nfl += STAPLE_FLOPS;
            }
          }
          staple2(stpl5, coef7, stpl3[pm[0]], U[pn[0]], sig, pn[0], coefL != 0 ? coefL : 0.0,
                  coefL != 0 ? stpl3[rho] : nullptr, stpl3[pm[1]], U[pn[1]], pn[1]);
          nfl += 18;
        } else {
          lincomb(stpl5, 0.0, coefL, stpl3[rho], 0.0, nullptr);
        }
        if (coefL != 0) nfl += 36;
        if (coefL != 0 || coef7 != 0) {
          side_force_mix(D[rho], Lin{mid[sig], 1.0, nullptr}, U[rho], stpl5, sig, rho, mid3,
                         Mix{coefL != 0 ? mid5[rho] : nullptr, coefL, coef7, coef5, mid[sig], m5set[rho]});
          if (coefL != 0) m5set[rho] = true;
          nfl += SIDE_FORCE_FLOPS;
          if (coefL != 0) nfl += 36;
        } else {
          lincomb(mid3, 0.0, coef5, mid[sig], 0.0, nullptr);
        }
        nfl += 18 + 36;
        for (int mu = 0; mu < 4; mu++) {
          if (mu == sig || mu == rho) continue;
          for (int nu = 0; nu < 4; nu++) {
            if (nu == sig || nu == rho || nu == mu) continue;
            side_force(D[mu], Lin{mid3, 1.0, nullptr}, U[mu], stpl3[nu], sig, mu, mid5[nu], m5set[nu]);
            m5set[nu] = true;
            nfl += SIDE_FORCE_FLOPS;
          }
        }
      }
      for (int mu = 0; mu < 4; mu++)
        if (mu != sig && !m5set[mu]) lincomb(mid5[mu], 0.0, 0.0, nullptr, 0.0, nullptr);
    }
    if (have3) {
      // the last one adds its staple sum (mid3) and coef1 mid to D[sig]
      int k = 0;
      for (int mu = 0; mu < 4; mu++) {
        if (mu == sig) continue;
        nfl += have5 ? 36 : 18;
        const Lin bot{mid[sig], coef3, have5 ? mid5[mu] : nullptr};
        if (++k < 3)
          side_force(D[mu], bot, U[mu], U[sig], sig, mu, mid3, k > 1);
        else
          side_force_fin(D[mu], bot, U[mu], U[sig], sig, mu, mid3, D[sig], coef1,
                         coef1 != 0 ? mid[sig] : nullptr);
        nfl += SIDE_FORCE_FLOPS;
      }
      nfl += 18;
    } else if (coef1 != 0) {
      lincomb(D[sig], 1.0, 0.0, nullptr, coef1, mid[sig]);
    }
    if (coef1 != 0) nfl += 36;
  }
  if (coefL != 0) {
    // Lepage fix-up (:317-339)
    const Ptr4 Up = {{U[0], U[1], U[2], U[3]}}, Mp = {{mid[0], mid[1], mid[2], mid[3]}};
    for (int mu = 0; mu < 4; mu++) {
      hipLaunchKernelGGL(lepage_acc_kernel, gx_, b_, 0, st_, g_, sq_, V_, mid3, Up, Mp, mu);
      hipLaunchKernelGGL(lepage_apply_kernel, gx_, b_, 0, st_, g_, sq_, V_, D[mu], (const double2 *)U[mu],
                         (const double2 *)mid3, -coefL, mu);
    }
    nfl += 4 * (3 * (2 * 198 + 3 * 18) + 198 + 216 + 36);
  }
  // Naik part (lib/asqtad_force_fnmat_p.c:395-445)
  if (coef[5] != 0 && midN) {
    for (int mu = 0; mu < 4; mu++)
      hipLaunchKernelGGL(naik_deriv_kernel, gx_, b_, 0, st_, g_, sq_, V_, D[mu], (const double2 *)U[mu],
                         (const double2 *)midN[mu], coef[5], mu);
    nfl += 4 * (4 * 198 + 216 + 18 + 36);
  }
}

}  // namespace

void asqtad_force_dev(const Lattice *lat, double *const U_raw[4], double *F_raw[4],
                      const double coef[6], const double *eps, const DevVec *const *xs, int n,
                      double *flops, hipStream_t st) {
  const bool naik = coef[5] != 0;
  const int nmid = naik ? 8 : 4;
  ForceRun r(lat, 4 + nmid + 4, eps, xs, n, st);
  double2 *U[4], *mid[8] = {}, *D[4];
  for (auto &u : U) u = r.take();
  for (int i = 0; i < nmid; i++) mid[i] = r.take();
  for (auto &d : D) d = r.take();
  for (int mu = 0; mu < 4; mu++) {
    r.to_soa(U_raw[mu], U[mu]);
    r.lincomb(D[mu], 0.0, 0.0, nullptr, 0.0, nullptr);
  }
  r.mids(0, n, nmid, -1.0, mid[0]);
  r.deriv(U, D, coef, mid, naik ? mid + 4 : nullptr);
  r.final(F_raw, U, D, -2.0);
  r.finish();
  if (flops) *flops = r.flops_per_site * (double)r.V();
}

// QOP_hisq_force_multi (lib/hisq_force_p.c:63-83 -> hisq_force_fnmat_p.c:7-231)
void hisq_force_dev(const Lattice *lat, double *const U_raw[4], double *const V_raw[4],
                    double *const W_raw[4], double *F_raw[4], const double fat7[6],
                    const double lev2[6], const double diff[6], int n_naiks, const double *eps_naik,
                    const double *eps, const DevVec *const *xs, const int *n_orders, double filter,
                    double *flops, int *nfilt, hipStream_t st) {
  int nterms = 0;
  for (int k = 0; k < n_naiks; k++) nterms += n_orders[k];
  ForceRun r(lat, 4 + 4 + 8 + 4 + 4, eps, xs, nterms, st);
  double2 *U[4], *W[4], *mid[8], *F1[4], *F2[4];
  for (auto &u : U) u = r.take();
  for (auto &w : W) w = r.take();
  for (auto &m : mid) m = r.take();
  for (auto &f : F1) f = r.take();
  for (auto &f : F2) f = r.take();
  DevBuf<int> cnt;
  dev_alloc(&cnt.p, sizeof(int));
  QOP_CHECK_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), st));
  for (int mu = 0; mu < 4; mu++) {
    r.to_soa(U_raw[mu], U[mu]);
    r.to_soa(W_raw[mu], W[mu]);
    r.lincomb(F2[mu], 0.0, 0.0, nullptr, 0.0, nullptr);
  }
  // level 2 per epsilon-Naik group (hisq_force_fnmat_p.c:56-118); group 0
  // takes the mids of all terms
  int shift = 0;
  for (int k = 0; k < n_naiks; k++) {
    const int c = k == 0 ? nterms : n_orders[k];
    r.mids(shift, c, 8, 1.0, mid[0]);
    for (int mu = 0; mu < 4; mu++) r.lincomb(F1[mu], 0.0, 0.0, nullptr, 0.0, nullptr);
    r.deriv(W, F1, k == 0 ? lev2 : diff, mid, mid + 4);
    const double cm = k == 0 ? 1.0 : eps_naik[k];
    for (int mu = 0; mu < 4; mu++) r.lincomb(F2[mu], 1.0, cm, F1[mu], 0.0, nullptr);
    r.flops_per_site += 4 * 36;
    shift += n_orders[k];
  }
  // U(3) projection derivative -> level-1 mids (:141-152)
  for (int mu = 0; mu < 4; mu++) r.u3_deriv(V_raw[mu], F2[mu], mid[mu], filter, cnt.p);
  r.flops_per_site += 4 * (198 + 81 * 16);
  // level 1: fat7 on U (:144-152), deriv = 2 x level 1, force += TA(U deriv^+)
  for (int mu = 0; mu < 4; mu++) r.lincomb(F1[mu], 0.0, 0.0, nullptr, 0.0, nullptr);
  double c7[6] = {fat7[0], fat7[1], fat7[2], fat7[3], fat7[4], 0.0};
  r.deriv(U, F1, c7, mid, nullptr);
  r.final(F_raw, U, F1, 2.0);
  r.flops_per_site += 4 * 36;
  int h = 0;
  QOP_CHECK_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
  r.finish();
  if (nfilt) *nfilt = h;
  if (flops) *flops = r.flops_per_site * (double)r.V();
}

// Asqtad fattening (lib/asqtad_dslash_p.c:391-633) of raw FP64 device links:
// fat = coef1' U + sum of staples (coef1' = coef1 - 6 lepage, :531), lng =
// naik U U U from gl; the order of accumulation is the reference's.
void fatten_dev(const Lattice *lat, const double *const gf[4], const double *const gl[4],
                const double c[6], double *fat_raw[4], double *lng_raw[4], hipStream_t st) {
  const long V = lat->V, n9 = 9 * V;
  if (V >= (1L << 27))
    throw std::runtime_error("fattening: lattice volume of 2^27 sites or more (32-bit offsets)");
  const Geom &g = lat->g;
  const double coef1 = c[0] - 6 * c[4], coef3 = c[1], coef5 = c[2], coef7 = c[3], coefL = c[4];
  const bool have5 = coef5 != 0 || coef7 != 0 || coefL != 0;
  const bool have3 = coef3 != 0 || have5;
  const bool naik = lng_raw != nullptr;
  const bool sep_long = naik && (gl[0] != gf[0] || gl[1] != gf[1] || gl[2] != gf[2] || gl[3] != gf[3]);
  const int nf = 4 + 1 + (have3 ? 2 : 0) + (naik ? 1 : 0) + (sep_long ? 4 : 0);
  DevBuf<double2> buf;
  dev_alloc(&buf.p, (size_t)nf * n9 * sizeof(double2));
  double2 *p = buf.p;
  auto take = [&] { double2 *r = p; p += n9; return r; };
  double2 *U[4], *UL[4], *F = take(), *S3 = nullptr, *S5 = nullptr, *Lk = nullptr;
  for (auto &u : U) u = take();
  if (have3) {
    S3 = take();
    S5 = take();
  }
  if (naik) Lk = take();
  for (int i = 0; i < 4; i++) UL[i] = sep_long ? take() : U[i];
  const Sched sq = make_sched(lat);
  const dim3 b(TPB), gs = grid(V), ge = grid(n9), gx(8 * sq.per);
  for (int mu = 0; mu < 4; mu++) {
    hipLaunchKernelGGL(to_soa_kernel, gs, b, 0, st, (const double2 *)gf[mu], U[mu], V);
    if (sep_long) hipLaunchKernelGGL(to_soa_kernel, gs, b, 0, st, (const double2 *)gl[mu], UL[mu], V);
  }
  auto add = [&](bool store, double2 *out, double cc, const double2 *I, const double2 *K, int mu, int nu) {
    if (store)
      hipLaunchKernelGGL(fat_staple_kernel<true>, gx, b, 0, st, g, sq, V, out, F, cc, I, K, mu, nu);
    else
      hipLaunchKernelGGL(fat_staple_kernel<false>, gx, b, 0, st, g, sq, V, out, F, cc, I, K, mu, nu);
  };
  for (int dir = 0; dir < 4; dir++) {
    hipLaunchKernelGGL(lincomb_kernel, ge, b, 0, st, F, 0.0, coef1, (const double2 *)U[dir], 0.0,
                       (const double2 *)nullptr, n9);
    if (have3) {
      for (int nu = 0; nu < 4; nu++) {
        if (nu == dir) continue;
        add(true, S3, coef3, U[dir], U[nu], dir, nu);
        if (coefL != 0) add(false, nullptr, coefL, S3, U[nu], dir, nu);
        if (coef5 != 0 || coef7 != 0)
          for (int rho = 0; rho < 4; rho++) {
            if (rho == dir || rho == nu) continue;
            add(coef7 != 0, S5, coef5, S3, U[rho], dir, rho);
            if (coef7 != 0)
              for (int sig = 0; sig < 4; sig++)
                if (sig != dir && sig != nu && sig != rho) add(false, nullptr, coef7, S5, U[sig], dir, sig);
          }
      }
    }
    hipLaunchKernelGGL(from_soa_kernel, gs, b, 0, st, (const double2 *)F, (double2 *)fat_raw[dir], V);
    if (naik) {
      hipLaunchKernelGGL(naik_link_kernel, gx, b, 0, st, g, sq, V, Lk, (const double2 *)UL[dir], c[5], dir);
      hipLaunchKernelGGL(from_soa_kernel, gs, b, 0, st, (const double2 *)Lk, (double2 *)lng_raw[dir], V);
    }
  }
  QOP_CHECK_HIP(hipGetLastError());
  QOP_CHECK_HIP(hipStreamSynchronize(st));
}

}  // namespace qop
