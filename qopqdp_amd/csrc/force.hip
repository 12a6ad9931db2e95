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

// raw [V][18] -> site-innermost [9][V]
__global__ __launch_bounds__(TPB) void to_soa_kernel(const double2 *raw, double2 *f, long V) {
  const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (s >= V) return;
#pragma unroll
  for (int e = 0; e < 9; e++) f[e * V + s] = raw[9 * s + e];
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

// QOP_get_mid (lib/asqtad_force_fnmat_p.c:4-44) with the factor -1 and the
// negation on odd sites (:491-503): mid_s(x) = -(-1)^p(x) sum_t eps_t
// x_t(x) x_t(x + dist_s)^+, dist_s = s-th direction (s < 4) or three times
// it (s >= 4).  Vectors are full-lattice [parity][colour][site] complex.
template <typename T2>
__global__ __launch_bounds__(TPB) void mid_kernel(Geom g, Sched q, long V, const T2 *const *xs, const double *eps, int n, int nmid,
                           double2 *mid) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const long Vh = g.Vh;
  const double sg = s >= Vh ? 1.0 : -1.0;
  auto vidx = [&](long i, int c) { return i + (i >= Vh ? 2 * Vh : 0) + c * Vh; };
  for (int m = 0; m < nmid; m++) {
    const unsigned j = nbr(g, x, m & 3, m < 4 ? 1 : 3, 0, 0);
    M3 acc = zero3();
    for (int t = 0; t < n; t++) {
      const T2 *v = xs[t];
      double2 a[3], b[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const T2 va = v[vidx(s, c)], vb = v[vidx(j, c)];
        a[c] = make_double2(va.x, va.y);
        b[c] = make_double2(vb.x, vb.y);
      }
      const double w = sg * eps[t];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double re = fma(a[r].x, b[c].x, a[r].y * b[c].y);
          const double im = fma(a[r].y, b[c].x, -a[r].x * b[c].y);
          acc.e[3 * r + c].x = fma(w, re, acc.e[3 * r + c].x);
          acc.e[3 * r + c].y = fma(w, im, acc.e[3 * r + c].y);
        }
    }
    st(mid + (long)m * 9 * V, V, s, acc);
  }
}

// staple (lib/asqtad_force_fnmat_p.c:93-127) with its scaling folded in:
// out(x) = [out(x) if acc] + sc [K(x) I(x+nu) K(x+mu)^+ + K(x-nu)^+ I(x-nu) K(x-nu+mu)]
//          + te E(x)
// I is the field along mu, K the link along nu.
template <bool ACC, bool HASE>
__global__ __launch_bounds__(TPB) void staple_kernel(Geom g, Sched q, long V, double2 *out, double sc, const double2 *I,
                              const double2 *K, int mu, int nu, double te, const double2 *E) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const unsigned xpnu = nbr(g, x, nu, 1, mu, 0), xpmu = nbr(g, x, mu, 1, nu, 0);
  const unsigned xmnu = nbr(g, x, nu, -1, mu, 0), xmnupmu = nbr(g, x, nu, -1, mu, 1);
  M3 r = ACC ? ld(out, V, s) : zero3();
  {
    const M3 t = mm<false, true>(ld(I, V, xpnu), ld(K, V, xpmu));
    axpy<false>(r, sc, mm<false, false>(ld(K, V, s), t));
  }
  {
    const M3 t = mm<true, false>(ld(K, V, xmnu), ld(I, V, xmnu));
    axpy<false>(r, sc, mm<false, false>(t, ld(K, V, xmnupmu)));
  }
  if constexpr (HASE) axpy<false>(r, te, ld(E, V, s));
  st(out, V, s, r);
}

// side_force (lib/asqtad_force_fnmat_p.c:129-199) at one site:
// force(x) += T(x) S(x+mu) B(x+nu)^+ + B(x) S(x+mu) T(x+nu)^+
//           + B(x-mu)^+ S(x-mu) T(x-mu+nu) + T(x-mu)^+ S(x-mu) B(x-mu+nu)
// stpl(x) = [stpl(x) if acc] + S(x) B(x+nu) S(x+mu)^+ + S(x-nu)^+ B(x-nu) S(x-nu+mu)
// with B = bot, S = side (the link along nu), T = top.
template <bool ACC, bool HASB>
__global__ __launch_bounds__(TPB) void side_force_kernel(Geom g, Sched q, long V, double2 *force, double2 *stpl,
                                                         Lin bot, const double2 *side,
                                                         const double2 *top, int mu, int nu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  const unsigned xpmu = nbr(g, x, mu, 1, nu, 0), xpnu = nbr(g, x, nu, 1, mu, 0);
  // SB() keeps the scheduler from hoisting every matrix load to the top:
  // with all 14 in flight a thread needs ~500 registers (one wave per SIMD).
#define SB() __builtin_amdgcn_sched_barrier(0)
  {
    const unsigned xmnu = nbr(g, x, nu, -1, mu, 0), xmnupmu = nbr(g, x, nu, -1, mu, 1);
    const M3 bs = mm<false, true>(ld<HASB>(bot, V, xpnu), ld(side, V, xpmu));  // B(x+nu) S(x+mu)^+
    SB();
    M3 sp = mm<false, false>(ld(side, V, s), bs);
    SB();
    const M3 t = mm<true, false>(ld(side, V, xmnu), ld<HASB>(bot, V, xmnu));
    SB();
    axpy<false>(sp, 1.0, mm<false, false>(t, ld(side, V, xmnupmu)));
    if constexpr (ACC) axpy<false>(sp, 1.0, ld(stpl, V, s));
    st(stpl, V, s, sp);
  }
  SB();
  const unsigned xmmu = nbr(g, x, mu, -1, nu, 0), xmmupnu = nbr(g, x, mu, -1, nu, 1);
  M3 f;
  {
    const M3 Smu = ld(side, V, xpmu);
    const M3 bs = mm<false, true>(ld<HASB>(bot, V, xpnu), Smu);
    SB();
    f = mm<false, true>(ld(top, V, s), bs);  // T S(x+mu) B(x+nu)^+
    SB();
    const M3 b2 = mm<false, false>(ld<HASB>(bot, V, s), Smu);  // B(x) S(x+mu)
    SB();
    axpy<false>(f, 1.0, mm<false, true>(b2, ld(top, V, xpnu)));
  }
  SB();
  {
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
#undef SB
  axpy<false>(f, 1.0, ld(force, V, s));
  st(force, V, s, f);
}

// After the Lepage/7-staple side_force (lib/asqtad_force_fnmat_p.c:268-277):
// mid5 += cl mid3 (if mid5), mid3 = c7 mid3 + c5 ms.
__global__ __launch_bounds__(TPB) void mix_kernel(double2 *mid5, double cl, double2 *mid3, double c7, double c5,
                           const double2 *ms, long n) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double2 t = mid3[k], m = ms[k];
  if (mid5) {
    double2 a = mid5[k];
    a.x = fma(cl, t.x, a.x);
    a.y = fma(cl, t.y, a.y);
    mid5[k] = a;
  }
  mid3[k] = make_double2(fma(c7, t.x, c5 * m.x), fma(c7, t.y, c5 * m.y));
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
  M3 d = ld(D, V, s);
  axpy<false>(d, cn, f);
  st(D, V, s, d);
}

// Lepage fix-up (lib/asqtad_force_fnmat_p.c:317-339), first half:
// acc(x) = sum_{nu != mu} [q_nu(x) + q_nu(x)^+], q_nu(x) = m_nu(x) U_nu(x)^+ - m_nu(x-nu)^+ U_nu(x-nu)
__global__ __launch_bounds__(TPB) void lepage_acc_kernel(Geom g, Sched q, long V, double2 *acc, const double2 *const *U,
                                  const double2 *const *m, int mu) {
  const int s0 = site_index(q, g.Vh);
  if (s0 < 0) return;
  const unsigned s = (unsigned)s0;
  const Site x = site_of(g, s);
  M3 r = zero3();
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    const unsigned b = nbr(g, x, nu, -1, 0, 0);
    M3 q = mm<false, true>(ld(m[nu], V, s), ld(U[nu], V, s));
    axpy<false>(q, -1.0, mm<true, false>(ld(m[nu], V, b), ld(U[nu], V, b)));
    axpy<false>(r, 1.0, q);
    axpy<true>(r, 1.0, q);
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
  M3 d = ld(D, V, s);
  axpy<false>(d, cf, t);
  st(D, V, s, d);
}

// force (raw layout) += -2 TA(U D^+), TA the traceless antihermitian part
// (lib/asqtad_force_fnmat_p.c:531-539)
__global__ __launch_bounds__(TPB) void force_kernel(long V, double2 *F, const double2 *U, const double2 *D) {
  const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (s >= V) return;
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
    double2 f = F[9 * s + e];
    f.x = fma(-2.0, t.e[e].x, f.x);
    f.y = fma(-2.0, t.e[e].y, f.y);
    F[9 * s + e] = f;
  }
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

void asqtad_force_dev(const Lattice *lat, double *const U_raw[4], double *F_raw[4],
                      const double coef[6], const double *eps, const DevVec *const *xs, int n,
                      double *flops, hipStream_t st) {
  const long V = lat->V;
  const Geom &g = lat->g;
  const long n9 = 9 * V;
  const bool naik = coef[5] != 0;
  double coef1 = coef[0], coef3 = coef[1], coef5 = coef[2], coef7 = coef[3], coefL = coef[4];
  coef1 -= 6 * coefL;
  const bool have5 = coef5 != 0 || coef7 != 0 || coefL != 0;
  const bool have3 = coef3 != 0 || have5;
  if (V >= (1L << 27))
    throw std::runtime_error("asqtad force: lattice volume of 2^27 sites or more (32-bit offsets)");
  const int prec = xs[0]->prec;
  for (int t = 0; t < n; t++)
    if (xs[t]->prec != prec || xs[t]->lat != lat || (prec != P64 && prec != P32))
      throw std::runtime_error("asqtad force: vectors must share the lattice and be FP64 or FP32");

  // site-innermost fields: U[4], mid[4 or 8], deriv[4], stpl3[4], mid5[4], stpl5, mid3
  const int nmid = naik ? 8 : 4;
  const int nf = 4 + nmid + 4 + (have5 ? 8 : 0) + 2;
  DevBuf<double2> buf;
  dev_alloc(&buf.p, (size_t)nf * n9 * sizeof(double2));
  double2 *p = buf.p;
  auto take = [&] { double2 *r = p; p += n9; return r; };
  double2 *U[4], *mid[8] = {}, *D[4], *stpl3[4] = {}, *mid5[4] = {};
  for (auto &u : U) u = take();
  for (int i = 0; i < nmid; i++) mid[i] = take();
  for (auto &d : D) d = take();
  if (have5) {
    for (auto &q : stpl3) q = take();
    for (auto &q : mid5) q = take();
  }
  double2 *stpl5 = take(), *mid3 = take();

  // vector pointers, eps and the U/mid pointer tables of the Lepage kernel
  const size_t ntab = (size_t)n * 2 + 8;
  DevBuf<double> tab;
  dev_alloc(&tab.p, ntab * sizeof(double));
  std::vector<double> htab(ntab);
  for (int t = 0; t < n; t++) {
    const void *d = xs[t]->data;
    memcpy(&htab[t], &d, sizeof d);
    htab[n + t] = eps[t];
  }
  for (int i = 0; i < 4; i++) {
    memcpy(&htab[2 * n + i], &U[i], sizeof U[i]);
    memcpy(&htab[2 * n + 4 + i], &mid[i], sizeof mid[i]);
  }
  QOP_CHECK_HIP(hipMemcpyAsync(tab.p, htab.data(), ntab * sizeof(double), hipMemcpyHostToDevice, st));
  const double *d_eps = tab.p + n;
  const double2 *const *d_U = (const double2 *const *)(tab.p + 2 * n);
  const double2 *const *d_mid = (const double2 *const *)(tab.p + 2 * n + 4);

  double nfl = 0;  // per site
  const Sched sq = make_sched(lat);
  const dim3 b(TPB), gs = grid(V), ge = grid(n9), gx(8 * sq.per);
  for (int mu = 0; mu < 4; mu++) {
    hipLaunchKernelGGL(to_soa_kernel, gs, b, 0, st, (const double2 *)U_raw[mu], U[mu], V);
    hipLaunchKernelGGL(lincomb_kernel, ge, b, 0, st, D[mu], 0.0, 0.0, (const double2 *)nullptr, 0.0,
                       (const double2 *)nullptr, n9);
  }
  if (prec == P64)
    hipLaunchKernelGGL(mid_kernel<double2>, gx, b, 0, st, g, sq, V, (const double2 *const *)tab.p, d_eps, n,
                       nmid, mid[0]);
  else
    hipLaunchKernelGGL(mid_kernel<float2>, gx, b, 0, st, g, sq, V, (const float2 *const *)tab.p, d_eps, n,
                       nmid, mid[0]);
  nfl += (54 + 36) * 4.0 * n * (naik ? 2 : 1);

  auto staple = [&](double2 *out, int acc, double sc, const double2 *I, const double2 *K, int mu,
                    int nu, double te, const double2 *E) {
    if (acc && E)
      hipLaunchKernelGGL((staple_kernel<true, true>), gx, b, 0, st, g, sq, V, out, sc, I, K, mu, nu, te, E);
    else if (acc)
      hipLaunchKernelGGL((staple_kernel<true, false>), gx, b, 0, st, g, sq, V, out, sc, I, K, mu, nu, te, E);
    else if (E)
      hipLaunchKernelGGL((staple_kernel<false, true>), gx, b, 0, st, g, sq, V, out, sc, I, K, mu, nu, te, E);
    else
      hipLaunchKernelGGL((staple_kernel<false, false>), gx, b, 0, st, g, sq, V, out, sc, I, K, mu, nu, te, E);
  };
  auto side_force = [&](double2 *force, Lin bot, const double2 *side, const double2 *top, int mu,
                        int nu, double2 *stpl, int acc) {
    if (acc && bot.b)
      hipLaunchKernelGGL((side_force_kernel<true, true>), gx, b, 0, st, g, sq, V, force, stpl, bot, side, top, mu, nu);
    else if (acc)
      hipLaunchKernelGGL((side_force_kernel<true, false>), gx, b, 0, st, g, sq, V, force, stpl, bot, side, top, mu, nu);
    else if (bot.b)
      hipLaunchKernelGGL((side_force_kernel<false, true>), gx, b, 0, st, g, sq, V, force, stpl, bot, side, top, mu, nu);
    else
      hipLaunchKernelGGL((side_force_kernel<false, false>), gx, b, 0, st, g, sq, V, force, stpl, bot, side, top, mu, nu);
  };
  auto lincomb = [&](double2 *o, double co, double ca, const double2 *A, double cb, const double2 *B) {
    hipLaunchKernelGGL(lincomb_kernel, ge, b, 0, st, o, co, ca, A, cb, B, n9);
  };

  // QOP_fat_deriv (lib/asqtad_force_fnmat_p.c:236-316)
  for (int sig = 0; sig < 4; sig++) {
    if (have5) {
      for (int mu = 0; mu < 4; mu++) {
        if (mu == sig) continue;
        staple(stpl3[mu], 0, 1.0, U[sig], U[mu], sig, mu, 0.0, nullptr);
        lincomb(mid5[mu], 0.0, 0.0, nullptr, 0.0, nullptr);
        nfl += STAPLE_FLOPS;
      }
      for (int rho = 0; rho < 4; rho++) {
        if (rho == sig) continue;
        // stpl5 = coef7 sum 7-staples + coefL stpl3[rho]
        if (coef7 != 0) {
          int first = 1;
          for (int mu = 0; mu < 4; mu++) {
            if (mu == sig || mu == rho) continue;
            for (int nu = 0; nu < 4; nu++) {
              if (nu == sig || nu == rho || nu == mu) continue;
              staple(stpl5, !first, coef7, stpl3[mu], U[nu], sig, nu, first && coefL != 0 ? coefL : 0.0,
                     first && coefL != 0 ? stpl3[rho] : nullptr);
              first = 0;
              nfl += STAPLE_FLOPS;
            }
          }
          nfl += 18;
        } else {
          lincomb(stpl5, 0.0, coefL, stpl3[rho], 0.0, nullptr);
        }
        if (coefL != 0) nfl += 36;
        if (coefL != 0 || coef7 != 0) {
          side_force(D[rho], Lin{mid[sig], 1.0, nullptr}, U[rho], stpl5, sig, rho, mid3, 0);
          nfl += SIDE_FORCE_FLOPS;
          if (coefL != 0) nfl += 36;
          hipLaunchKernelGGL(mix_kernel, ge, b, 0, st, coefL != 0 ? mid5[rho] : nullptr, coefL, mid3,
                             coef7, coef5, (const double2 *)mid[sig], n9);
        } else {
          lincomb(mid3, 0.0, coef5, mid[sig], 0.0, nullptr);
        }
        nfl += 18 + 36;
        for (int mu = 0; mu < 4; mu++) {
          if (mu == sig || mu == rho) continue;
          for (int nu = 0; nu < 4; nu++) {
            if (nu == sig || nu == rho || nu == mu) continue;
            side_force(D[mu], Lin{mid3, 1.0, nullptr}, U[mu], stpl3[nu], sig, mu, mid5[nu], 1);
            nfl += SIDE_FORCE_FLOPS;
          }
        }
      }
    }
    if (have3) {
      int first = 1;
      for (int mu = 0; mu < 4; mu++) {
        if (mu == sig) continue;
        nfl += have5 ? 36 : 18;
        side_force(D[mu], Lin{mid[sig], coef3, have5 ? mid5[mu] : nullptr}, U[mu], U[sig], sig, mu, mid3,
                   !first);
        first = 0;
        nfl += SIDE_FORCE_FLOPS;
      }
      nfl += 18;
    }
    if (coef1 != 0) nfl += 36;
    if (have3 || coef1 != 0)
      lincomb(D[sig], 1.0, have3 ? 1.0 : 0.0, have3 ? mid3 : nullptr, coef1, coef1 != 0 ? mid[sig] : nullptr);
  }
  if (coefL != 0) {
    for (int mu = 0; mu < 4; mu++) {
      hipLaunchKernelGGL(lepage_acc_kernel, gx, b, 0, st, g, sq, V, mid3, d_U, d_mid, mu);
      hipLaunchKernelGGL(lepage_apply_kernel, gx, b, 0, st, g, sq, V, D[mu], (const double2 *)U[mu],
                         (const double2 *)mid3, -coefL, mu);
    }
    nfl += 4 * (3 * (2 * 198 + 3 * 18) + 198 + 216 + 36);
  }
  // Naik part (lib/asqtad_force_fnmat_p.c:395-445)
  if (naik) {
    for (int mu = 0; mu < 4; mu++)
      hipLaunchKernelGGL(naik_deriv_kernel, gx, b, 0, st, g, sq, V, D[mu], (const double2 *)U[mu],
                         (const double2 *)mid[4 + mu], coef[5], mu);
    nfl += 4 * (4 * 198 + 216 + 18 + 36);
  }
  for (int mu = 0; mu < 4; mu++)
    hipLaunchKernelGGL(force_kernel, gs, b, 0, st, V, (double2 *)F_raw[mu], (const double2 *)U[mu],
                       (const double2 *)D[mu]);
  nfl += 4 * (198 + 24 + 18);
  QOP_CHECK_HIP(hipGetLastError());
  QOP_CHECK_HIP(hipStreamSynchronize(st));
  if (flops) *flops = nfl * (double)V;
}

}  // namespace qop
