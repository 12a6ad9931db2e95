/*
 * CPU oracle for the asqtad staggered hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 (OpenMP) restatement of the reference algorithms.  Only
 * tests/, the smoke check and the cpu_baseline leg of bench.py load it, and
 * only as the checker; the product library never links it.
 *
 * Function -> reference file:line it follows -> what pins it (tests/):
 *   or_raw_index/coords   lib/layout_user.c:237-258        round trip + parity tests
 *   or_rephase_G          lib/common_p.c:464-526           closed form per site; free-field spectrum
 *   or_make_imp_links     lib/asqtad_dslash_p.c:391-633    free-field weights 9/8, -1/24; gauge covariance;
 *                                                          one-link-only and naik-only closed forms
 *   or_links_create       lib/asqtad_dslash_p.c:79-223     x0.5 and double store pinned by check_op agreement
 *   or_dslash             lib/asqtad_dslash_p.c:997-1092   independent check_op restatement
 *                                                          (examples/asqtad_test.c:37-81), brute-force dense
 *                                                          matrix on 4^4, anti-hermiticity, free-field spectrum,
 *                                                          gauge covariance
 *   or_dslash_dir         lib/asqtad_dslash_p.c:1140-1377  sum over directions and fb of the one-direction
 *                                                          transports = the hop and = check_op (m = 0)
 *   or_links_set_single   lib/pcstubs/asqtad_invert_f3.c   stored values FP32-representable; FP32 solves
 *                                                          agree with FP64 to FP32 accuracy
 *   or_ddagd_norm2        lib/asqtad_dslash_p.c:1121-1138  <p,Ap> identity, free-field spectrum
 *   or_invert (EVEN/ODD)  lib/asqtad_invert_p.c:246-647    free-field exact solution; check_resids identity
 *   or_invert (EVENODD)   lib/asqtad_invert_p.c:649-851    full-D residual ||in - D out|| and free field
 *   invert_cg / cgms      lib/invert_cg_p.c:1-402          dense-solve agreement on 4^4; shifted solutions
 *   or_invert_multi       lib/asqtad_invert_p.c:853-1059   per-mass residuals; multi vs single mass agreement
 *   or_relnorm2           lib/invert_utilities_p.c:7-40    closed form on constructed fields
 *   or_plaquette          examples/test_common.c:171-228   unit gauge -> 1; gauge invariance
 */
#include "qop_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int verbose = 0;
int or_set_verbose(int v) { int o = verbose; verbose = v; return o; }
int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */
long or_volume(const int L[4]) { return (long)L[0] * L[1] * L[2] * L[3]; }

long or_raw_index(const int L[4], const int x[4]) {
  long l = ((long)(x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3];
  int p = (x[0] + x[1] + x[2] + x[3]) & 1;
  return p ? (l + or_volume(L)) / 2 : l / 2;
}

static void lex_coords(const int L[4], long l, int x[4]) {
  x[3] = (int)(l % L[3]); l /= L[3];
  x[2] = (int)(l % L[2]); l /= L[2];
  x[1] = (int)(l % L[1]); l /= L[1];
  x[0] = (int)l;
}

void or_raw_coords(const int L[4], long idx, int x[4]) {
  long V = or_volume(L), Vh = V / 2;
  int p = idx >= Vh;
  long h = p ? idx - Vh : idx;
  lex_coords(L, 2 * h, x);
  if (((x[0] + x[1] + x[2] + x[3]) & 1) != p) lex_coords(L, 2 * h + 1, x);
}

static long shift_index(const int L[4], const int x[4], int mu, int d) {
  int y[4] = {x[0], x[1], x[2], x[3]};
  y[mu] = ((y[mu] + d) % L[mu] + L[mu]) % L[mu];
  return or_raw_index(L, y);
}

/* ------------------------------------------------------------------ */
/* 3x3 complex helpers (m[i][j][re/im] stored as 18 doubles)           */
/* ------------------------------------------------------------------ */
#define RE(m, i, j) (m)[6 * (i) + 2 * (j)]
#define IM(m, i, j) (m)[6 * (i) + 2 * (j) + 1]

static void m_zero(double *c) { memset(c, 0, 18 * sizeof(double)); }
/* c = a*b */
static void m_mul(double *c, const double *a, const double *b) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, i, k) * RE(b, k, j) - IM(a, i, k) * IM(b, k, j);
        im += RE(a, i, k) * IM(b, k, j) + IM(a, i, k) * RE(b, k, j);
      }
      RE(t, i, j) = re; IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}
/* c = a*b^dagger */
static void m_mul_na(double *c, const double *a, const double *b) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, i, k) * RE(b, j, k) + IM(a, i, k) * IM(b, j, k);
        im += -RE(a, i, k) * IM(b, j, k) + IM(a, i, k) * RE(b, j, k);
      }
      RE(t, i, j) = re; IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}
/* c = a^dagger*b */
static void m_mul_an(double *c, const double *a, const double *b) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, k, i) * RE(b, k, j) + IM(a, k, i) * IM(b, k, j);
        im += RE(a, k, i) * IM(b, k, j) - IM(a, k, i) * RE(b, k, j);
      }
      RE(t, i, j) = re; IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}
/* c += s*a */
static void m_axpy(double *c, double s, const double *a) {
  for (int i = 0; i < 18; i++) c[i] += s * a[i];
}
/* out += M v */
static void mv_acc(double *out, const double *m, const double *v) {
  for (int i = 0; i < 3; i++) {
    double re = 0, im = 0;
    for (int j = 0; j < 3; j++) {
      re += RE(m, i, j) * v[2 * j] - IM(m, i, j) * v[2 * j + 1];
      im += RE(m, i, j) * v[2 * j + 1] + IM(m, i, j) * v[2 * j];
    }
    out[2 * i] += re; out[2 * i + 1] += im;
  }
}
/* out += M^dagger v */
static void mav_acc(double *out, const double *m, const double *v) {
  for (int i = 0; i < 3; i++) {
    double re = 0, im = 0;
    for (int j = 0; j < 3; j++) {
      re += RE(m, j, i) * v[2 * j] + IM(m, j, i) * v[2 * j + 1];
      im += RE(m, j, i) * v[2 * j + 1] - IM(m, j, i) * v[2 * j];
    }
    out[2 * i] += re; out[2 * i + 1] += im;
  }
}

/* ------------------------------------------------------------------ */
/* rephase_G (lib/common_p.c:464-526)                                  */
/* ------------------------------------------------------------------ */
void or_rephase_G(const int L[4], double *U[4], const int r0[4],
                  const double bc[8], const int signmask[4]) {
  long V = or_volume(L);
  for (int mu = 0; mu < 4; mu++) {
    int bcdir = (bc && (bc[2 * mu] != 1.0 || bc[2 * mu + 1] != 0.0));
    int bits = signmask ? signmask[mu] : 0;
#pragma omp parallel for
    for (long i = 0; i < V; i++) {
      int x[4];
      or_raw_coords(L, i, x);
      double *m = U[mu] + 18 * i;
      if (bcdir) {
        int rs = (x[mu] + L[mu] - r0[mu]) % L[mu];
        if (rs == L[mu] - 1) {
          double pr = bc[2 * mu], pi = bc[2 * mu + 1];
          for (int e = 0; e < 9; e++) {
            double a = m[2 * e], b = m[2 * e + 1];
            m[2 * e] = pr * a - pi * b;
            m[2 * e + 1] = pr * b + pi * a;
          }
        }
      }
      if (bits) {
        /* the sign uses absolute coordinates (common_p.c:476-482) */
        int s = 0;
        for (int j = 0; j < 4; j++) s += ((bits >> j) & 1) * x[j];
        if (s & 1) for (int e = 0; e < 18; e++) m[e] = -m[e];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* plaquette (examples/test_common.c:171-228)                          */
/* ------------------------------------------------------------------ */
double or_plaquette(const int L[4], double *const U[4]) {
  long V = or_volume(L);
  double sum = 0;
  for (int mu = 0; mu < 3; mu++)
    for (int nu = mu + 1; nu < 4; nu++) {
#pragma omp parallel for reduction(+ : sum)
      for (long i = 0; i < V; i++) {
        int x[4];
        or_raw_coords(L, i, x);
        long xpmu = shift_index(L, x, mu, 1), xpnu = shift_index(L, x, nu, 1);
        double t3[18], t4[18];
        m_mul_an(t3, U[nu] + 18 * i, U[mu] + 18 * i);
        m_mul(t4, t3, U[nu] + 18 * xpmu);
        const double *t2 = U[mu] + 18 * xpnu;
        double tr = 0;
        for (int e = 0; e < 18; e++) tr += t2[e] * t4[e]; /* Re tr(t2^dag t4) */
        sum += tr;
      }
    }
  return sum / (6.0 * V * 3.0);
}

/* ------------------------------------------------------------------ */
/* fattening (lib/asqtad_dslash_p.c:391-633)                           */
/* ------------------------------------------------------------------ */
/* Generalised staple of `link` in the (mu,nu) plane (compute_gen_staple,
 * asqtad_dslash_p.c:391-455): upper(x) = U_nu(x) link(x+nu) U_nu(x+mu)^+,
 * lower(x) = U_nu(x-nu)^+ link(x-nu) U_nu(x-nu+mu).  If staple != NULL it is
 * saved as upper+lower and fl += coef*staple, else fl += coef*upper, then
 * fl += coef*lower (same accumulation order as the reference). */
static void gen_staple(const int L[4], double *staple, int mu, int nu,
                       const double *link, double coef, double *const gf[4],
                       double *fl) {
  long V = or_volume(L);
#pragma omp parallel for
  for (long i = 0; i < V; i++) {
    int x[4], y[4];
    or_raw_coords(L, i, x);
    long xpnu = shift_index(L, x, nu, 1), xpmu = shift_index(L, x, mu, 1);
    long xmnu = shift_index(L, x, nu, -1);
    or_raw_coords(L, xmnu, y);
    long xmnupmu = shift_index(L, y, mu, 1);
    double t1[18], up[18], lo[18];
    m_mul_na(t1, link + 18 * xpnu, gf[nu] + 18 * xpmu);
    m_mul(up, gf[nu] + 18 * i, t1);
    m_mul_an(t1, gf[nu] + 18 * xmnu, link + 18 * xmnu);
    m_mul(lo, t1, gf[nu] + 18 * xmnupmu);
    double *f = fl + 18 * i;
    if (staple) {
      double *s = staple + 18 * i;
      for (int e = 0; e < 18; e++) s[e] = up[e] + lo[e];
      m_axpy(f, coef, s);
    } else {
      m_axpy(f, coef, up);
      m_axpy(f, coef, lo);
    }
  }
}

void or_make_imp_links(const int L[4], const or_asqtad_coeffs_t *c,
                       double *const gf[4], double *const gfLong[4],
                       double *fat[4], double *lng[4]) {
  long V = or_volume(L);
  double coef1 = c->one_link, coef3 = c->three_staple, coef5 = c->five_staple;
  double coef7 = c->seven_staple, coefL = c->lepage;
  coef1 -= 6 * coefL; /* asqtad_dslash_p.c:531 */
  int have5 = coef5 != 0 || coef7 != 0 || coefL != 0;
  int have3 = coef3 != 0 || have5;
  double *staple = malloc(V * 18 * sizeof(double));
  double *temp = malloc(V * 18 * sizeof(double));
  for (int dir = 0; dir < 4; dir++) {
    for (long i = 0; i < 18 * V; i++) fat[dir][i] = coef1 * gf[dir][i];
    if (!have3) continue;
    for (int nu = 0; nu < 4; nu++) {
      if (nu == dir) continue;
      gen_staple(L, staple, dir, nu, gf[dir], coef3, gf, fat[dir]);
      if (coefL != 0) gen_staple(L, NULL, dir, nu, staple, coefL, gf, fat[dir]);
      if (coef5 != 0 || coef7 != 0) {
        for (int rho = 0; rho < 4; rho++) {
          if (rho == dir || rho == nu) continue;
          gen_staple(L, temp, dir, rho, staple, coef5, gf, fat[dir]);
          if (coef7 != 0)
            for (int sig = 0; sig < 4; sig++)
              if (sig != dir && sig != nu && sig != rho)
                gen_staple(L, NULL, dir, sig, temp, coef7, gf, fat[dir]);
        }
      }
    }
  }
  if (lng) { /* asqtad_dslash_p.c:594-604: L = naik * U(x) [U(x+mu) U(x+2mu)] */
    for (int dir = 0; dir < 4; dir++) {
#pragma omp parallel for
      for (long i = 0; i < V; i++) {
        int x[4];
        or_raw_coords(L, i, x);
        long x1 = shift_index(L, x, dir, 1), x2 = shift_index(L, x, dir, 2);
        double t[18], o[18];
        m_mul(t, gfLong[dir] + 18 * x1, gfLong[dir] + 18 * x2);
        m_mul(o, gfLong[dir] + 18 * i, t);
        for (int e = 0; e < 18; e++) lng[dir][18 * i + e] = c->naik * o[e];
      }
    }
  }
  free(staple);
  free(temp);
}

/* ------------------------------------------------------------------ */
/* link object and dslash                                              */
/* ------------------------------------------------------------------ */
struct or_links_s {
  int L[4];
  long V, Vh;
  int nlinks;
  double *dbl[16]; /* W_k(x), k = 4mu+{0:fat fwd, 1:long fwd, 2:fat bck, 3:long bck} */
  int disp[16], dir[16];
  int single; /* FP32 storage of links and vectors (the F3 build), or_links_set_single */
};

/* A value as the link object stores it: rounded to FP32 in single mode. */
static inline double ST(const or_links_t *l, double v) { return l->single ? (double)(float)v : v; }

/* Single-precision restatement (lib/pcstubs/asqtad_invert_f3.c instantiates
 * the same drivers with QOP_Precision 'F'): links and every vector the
 * operator and the solvers store are rounded to FP32; sums, dot products
 * and the solver scalars stay FP64 (QLA_D_Real, invert_cg_p.c:13-16). */
void or_links_set_single(or_links_t *l, int on) {
  l->single = on;
  if (!on) return;
  for (int k = 0; k < l->nlinks; k++)
    for (long i = 0; i < 18 * l->V; i++) l->dbl[k][i] = (double)(float)l->dbl[k][i];
}

or_links_t *or_links_create(const int L[4], double *const fat[4],
                            double *const lng[4]) {
  or_links_t *l = calloc(1, sizeof *l);
  memcpy(l->L, L, sizeof l->L);
  l->V = or_volume(L);
  l->Vh = l->V / 2;
  l->nlinks = lng ? 16 : 8;
  int nl2 = l->nlinks / 2;
  for (int k = 0; k < l->nlinks; k++) l->dbl[k] = malloc(18 * l->V * sizeof(double));
  for (int mu = 0; mu < 4; mu++) {
    /* fwd links scaled by 1/2 (asqtad_dslash_p.c:201-206) and back links
     * bck(x) = -fwd(x - s)^+ (asqtad_dslash_p.c:123-127) */
    int nper = nl2 / 4; /* 2 with long links, 1 without */
    for (int j = 0; j < nper; j++) {
      const double *src = j == 0 ? fat[mu] : lng[mu];
      int d = j == 0 ? 1 : 3;
      int kf = nl2 == 8 ? 4 * mu + j : 2 * mu;
      int kb = nl2 == 8 ? 4 * mu + 2 + j : 2 * mu + 1;
      l->disp[kf] = d; l->dir[kf] = mu;
      l->disp[kb] = -d; l->dir[kb] = mu;
      double *wf = l->dbl[kf], *wb = l->dbl[kb];
#pragma omp parallel for
      for (long i = 0; i < l->V; i++)
        for (int e = 0; e < 18; e++) wf[18 * i + e] = 0.5 * src[18 * i + e];
#pragma omp parallel for
      for (long i = 0; i < l->V; i++) {
        int x[4];
        or_raw_coords(L, i, x);
        long xm = shift_index(L, x, mu, -d);
        const double *m = wf + 18 * xm;
        double *o = wb + 18 * i;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            RE(o, a, b) = -RE(m, b, a);
            IM(o, a, b) = IM(m, b, a);
          }
      }
    }
  }
  return l;
}

void or_links_destroy(or_links_t *l) {
  if (!l) return;
  for (int k = 0; k < l->nlinks; k++) free(l->dbl[k]);
  free(l);
}

static void sub_range(const or_links_t *l, int eo, long *b, long *e) {
  *b = eo == OR_ODD ? l->Vh : 0;
  *e = eo == OR_EVEN ? l->Vh : l->V;
}

/* asqtad_dslash1 (asqtad_dslash_p.c:1246-1279): dest(x) = sum_k W_k(x) src(x+d_k)
 * on the sites of `eo`.  The source is first copied (asqtad_hop macro,
 * :997-1017), so dest may alias src. */
static void hop(or_links_t *l, double *dest, const double *src, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
  double *tsrc = malloc(6 * l->V * sizeof(double));
  memcpy(tsrc, src, 6 * l->V * sizeof(double));
#pragma omp parallel for
  for (long i = b; i < e; i++) {
    int x[4];
    or_raw_coords(l->L, i, x);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < l->nlinks; k++) {
      long j = shift_index(l->L, x, l->dir[k], l->disp[k]);
      mv_acc(acc, l->dbl[k] + 18 * i, tsrc + 6 * j);
    }
    for (int c = 0; c < 6; c++) dest[6 * i + c] = ST(l, acc[c]);
  }
  free(tsrc);
}

/* out = mass*in (+ sum) on subset (mass_term, asqtad_dslash_p.c:1019-1023) */
static void mass_term(or_links_t *l, double *out, const double *in, double mass,
                      int add, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
  for (long i = 6 * b; i < 6 * e; i++) out[i] = ST(l, mass * in[i] + (add ? out[i] : 0.0));
}

void or_dslash(or_links_t *l, double mass, double *out, const double *in,
               int eo_out, int eo_in) {
  if (eo_in == eo_out) {
    if (eo_out == OR_EVENODD) {
      hop(l, out, in, OR_EVENODD);
      mass_term(l, out, in, mass, 1, OR_EVENODD);
    } else {
      mass_term(l, out, in, mass, 0, eo_out);
    }
    return;
  }
  if (eo_out == OR_EVEN || eo_out == OR_EVENODD) {
    if (eo_in == OR_ODD) hop(l, out, in, OR_EVEN);
    else if (eo_in == OR_EVEN) mass_term(l, out, in, mass, 0, OR_EVEN);
    else { hop(l, out, in, OR_EVEN); mass_term(l, out, in, mass, 1, OR_EVEN); }
  }
  if (eo_out == OR_ODD || eo_out == OR_EVENODD) {
    if (eo_in == OR_EVEN) hop(l, out, in, OR_ODD);
    else if (eo_in == OR_ODD) mass_term(l, out, in, mass, 0, OR_ODD);
    else { hop(l, out, in, OR_ODD); mass_term(l, out, in, mass, 1, OR_ODD); }
  }
}

/* QOP_asqtad_dslash_dir (asqtad_dslash_p.c:1140-1191, style 1 :1338-1377):
 * out(x) += r_j W_k(x) in(x + d_k) on the sites of eo_out for the fat and
 * long link of direction dir, forward (fb > 0) or backward; bsign multiplies
 * the backward terms: +1 with the stored -U^+ of the double-stored styles,
 * -1 for style 0, whose backward term enters as +U^+ (:1280-1336).  Without
 * long links only the fat term (the reference indexes past its links there). */
void or_dslash_dir(or_links_t *l, int dir, int fb, double wtfat, double wtlong, double *out,
                   const double *in, int eo_out, double bsign) {
  long b, e;
  sub_range(l, eo_out, &b, &e);
  double *tsrc = malloc(6 * l->V * sizeof(double));
  memcpy(tsrc, in, 6 * l->V * sizeof(double));
  const int nper = l->nlinks == 16 ? 2 : 1;
#pragma omp parallel for
  for (long i = b; i < e; i++) {
    int x[4];
    or_raw_coords(l->L, i, x);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nper; j++) {
      int k = l->nlinks == 16 ? 4 * dir + (fb > 0 ? 0 : 2) + j : 2 * dir + (fb > 0 ? 0 : 1);
      long jn = shift_index(l->L, x, dir, l->disp[k]);
      double w = (j ? wtlong : wtfat) * (fb > 0 ? 1.0 : bsign), t[6] = {0, 0, 0, 0, 0, 0};
      mv_acc(t, l->dbl[k] + 18 * i, tsrc + 6 * jn);
      for (int c = 0; c < 6; c++) acc[c] += w * t[c];
    }
    for (int c = 0; c < 6; c++) out[6 * i + c] = ST(l, out[6 * i + c] + acc[c]);
  }
  free(tsrc);
}

/* check_op restatement (examples/asqtad_test.c:37-81):
 * out = 1/2 [ 2m in + sum_mu ( F(x) in(x+mu) - F(x-mu)^+ in(x-mu)
 *                            + L(x) in(x+3mu) - L(x-3mu)^+ in(x-3mu) ) ]
 * using the user-normalised links directly (no x0.5 pre-scaling and no
 * double storage), so it is an independent second path. */
void or_dslash_check_op(const int L[4], double *const fat[4],
                        double *const lng[4], double mass, double *out,
                        const double *in) {
  long V = or_volume(L);
#pragma omp parallel for
  for (long i = 0; i < V; i++) {
    int x[4];
    or_raw_coords(L, i, x);
    double chk[6], neg[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 6; c++) chk[c] = 2 * mass * in[6 * i + c];
    for (int mu = 0; mu < 4; mu++) {
      long xp = shift_index(L, x, mu, 1), xm = shift_index(L, x, mu, -1);
      mv_acc(chk, fat[mu] + 18 * i, in + 6 * xp);
      mav_acc(neg, fat[mu] + 18 * xm, in + 6 * xm);
      if (lng) {
        long xp3 = shift_index(L, x, mu, 3), xm3 = shift_index(L, x, mu, -3);
        mv_acc(chk, lng[mu] + 18 * i, in + 6 * xp3);
        mav_acc(neg, lng[mu] + 18 * xm3, in + 6 * xm3);
      }
    }
    for (int c = 0; c < 6; c++) out[6 * i + c] = 0.5 * (chk[c] - neg[c]);
  }
}

/* ------------------------------------------------------------------ */
/* vector helpers on a subset                                          */
/* ------------------------------------------------------------------ */
static double norm2(or_links_t *l, const double *v, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
  double s = 0;
#pragma omp parallel for reduction(+ : s)
  for (long i = 6 * b; i < 6 * e; i++) s += v[i] * v[i];
  return s;
}
/* sum conj(a) b (QDP_c_eq_V_dot_V) */
static void cdot(or_links_t *l, const double *a, const double *b_, int eo,
                 double *re, double *im) {
  long b, e;
  sub_range(l, eo, &b, &e);
  double sr = 0, si = 0;
#pragma omp parallel for reduction(+ : sr, si)
  for (long i = b; i < e; i++)
    for (int c = 0; c < 3; c++) {
      double ar = a[6 * i + 2 * c], ai = a[6 * i + 2 * c + 1];
      double br = b_[6 * i + 2 * c], bi = b_[6 * i + 2 * c + 1];
      sr += ar * br + ai * bi;
      si += ar * bi - ai * br;
    }
  *re = sr; *im = si;
}
static double redot(or_links_t *l, const double *a, const double *b_, int eo) {
  double r, i;
  cdot(l, a, b_, eo, &r, &i);
  return r;
}
static double *vnew(or_links_t *l) { return calloc(6 * l->V, sizeof(double)); }
static void vcopy(or_links_t *l, double *d, const double *s, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
  memcpy(d + 6 * b, s + 6 * b, 6 * (e - b) * sizeof(double));
}
static void vzero(or_links_t *l, double *d, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
  memset(d + 6 * b, 0, 6 * (e - b) * sizeof(double));
}
/* d = a*x + y (real a) */
static void vaxpy(or_links_t *l, double *d, double a, const double *x,
                  const double *y, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
#pragma omp parallel for
  for (long i = 6 * b; i < 6 * e; i++) d[i] = ST(l, a * x[i] + y[i]);
}
/* d += (cr + i ci) * x */
static void vcaxpy(or_links_t *l, double *d, double cr, double ci,
                   const double *x, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
#pragma omp parallel for
  for (long i = b; i < e; i++)
    for (int c = 0; c < 3; c++) {
      double xr = x[6 * i + 2 * c], xi = x[6 * i + 2 * c + 1];
      d[6 * i + 2 * c] = ST(l, d[6 * i + 2 * c] + (cr * xr - ci * xi));
      d[6 * i + 2 * c + 1] = ST(l, d[6 * i + 2 * c + 1] + (cr * xi + ci * xr));
    }
}

static int oppsub(int eo) { return (4 - eo) % 3; } /* qop_internal.h:64 */

/* ddagd_norm2 (asqtad_dslash_p.c:1121-1138) with its own temporary */
double or_ddagd_norm2(or_links_t *l, double mass, double *out,
                      const double *in, int eo) {
  double m2 = mass * mass;
  double *tmp = vnew(l);
  or_dslash(l, mass, tmp, in, oppsub(eo), eo);
  or_dslash(l, mass, out, tmp, eo, oppsub(eo));
  long b, e;
  sub_range(l, eo, &b, &e);
  for (long i = 6 * b; i < 6 * e; i++) out[i] = ST(l, m2 * in[i] - out[i]);
  double in2 = norm2(l, in, eo), Ain2 = norm2(l, tmp, oppsub(eo));
  free(tmp);
  return Ain2 + m2 * in2;
}

/* relnorm2 (invert_utilities_p.c:7-40) */
double or_relnorm2(or_links_t *l, const double *r, const double *x, int eo) {
  long b, e;
  sub_range(l, eo, &b, &e);
  double g = 0;
  for (long i = b; i < e; i++) {
    double x2 = 0, r2 = 0;
    for (int c = 0; c < 6; c++) { x2 += x[6 * i + c] * x[6 * i + c]; r2 += r[6 * i + c] * r[6 * i + c]; }
    g += x2 == 0.0 ? 1.0 : r2 / x2;
  }
  return g / (double)(e - b);
}

/* ------------------------------------------------------------------ */
/* Krylov solvers (lib/invert_cg_p.c)                                  */
/* ------------------------------------------------------------------ */
typedef struct {
  or_links_t *l;
  double mass;
  int eo;
} linop_ctx;

static double lin_norm2(linop_ctx *c, double *out, const double *in) {
  return or_ddagd_norm2(c->l, c->mass, out, in, c->eo);
}
static void lin_plain(linop_ctx *c, double *out, const double *in) {
  or_ddagd_norm2(c->l, c->mass, out, in, c->eo); /* ddagd: same output */
}

/* regular CG (invert_cg_p.c:1-128) */
static void invert_cg(linop_ctx *lc, or_invert_arg_t *ia, or_resid_arg_t *ra,
                      double *out, const double *in, double *p) {
  or_links_t *l = lc->l;
  int eo = lc->eo;
  double a, b, rsq, oldrsq, pkp, relnorm2, insq, rsqstop;
  int iteration = 0, total = 0, nrestart = -1;
  int restart = ia->restart, max_iter = ia->max_iter, max_restarts = ia->max_restarts;
  if (max_restarts < 0) max_restarts = 5;
  ra->final_rsq = 0; ra->final_rel = 0; ra->final_iter = 0; ra->final_restart = 0;
  insq = norm2(l, in, eo);
  if (insq == 0.) { vzero(l, out, eo); return; }
  double *r = vnew(l), *Mp = vnew(l);
  rsqstop = ra->rsqmin * insq;
  rsq = 0; relnorm2 = 1.; oldrsq = rsq;
  while (1) {
    if (total == 0 || iteration >= restart || total >= max_iter || rsq < rsqstop ||
        relnorm2 < ra->relmin) {
      if (total >= max_iter || nrestart >= max_restarts) break;
      nrestart++;
      vcopy(l, p, out, eo);
      lin_plain(lc, Mp, p);
      iteration = 1;
      total++;
      vaxpy(l, r, -1.0, Mp, in, eo);
      rsq = norm2(l, r, eo);
      if (ra->relmin > 0) relnorm2 = or_relnorm2(l, r, out, eo);
      if (verbose) printf("CG: (re)start: iter %i rsq = %g rel = %g\n", total, rsq, relnorm2);
      if (rsq < rsqstop || relnorm2 < ra->relmin || total >= max_iter) break;
      vcopy(l, p, r, eo);
    } else {
      b = rsq / oldrsq;
      vaxpy(l, p, b, p, r, eo);
    }
    oldrsq = rsq;
    lin_plain(lc, Mp, p);
    iteration++;
    total++;
    pkp = redot(l, p, Mp, eo);
    if (pkp <= 0) break;
    a = rsq / pkp;
    vaxpy(l, out, a, p, out, eo);
    vaxpy(l, r, -a, Mp, r, eo);
    rsq = norm2(l, r, eo);
    if (ra->relmin > 0) relnorm2 = or_relnorm2(l, r, out, eo);
  }
  free(r); free(Mp);
  ra->final_rsq = rsq / insq;
  ra->final_rel = relnorm2;
  ra->final_iter = total;
  ra->final_restart = nrestart;
}

/* multi-shift CG (invert_cg_p.c:130-402) */
static void invert_cgms(linop_ctx *lc, or_invert_arg_t *ia, or_resid_arg_t **ra,
                        const double *shifts, int ns, double **out,
                        const double *in, double *p) {
  or_links_t *l = lc->l;
  int eo = lc->eo;
  double a[ns], b[ns], bo[ns], z[ns], zo[ns], zn[ns];
  double rsq, oldrsq, pkp, relnorm2, insq;
  int iteration = 0;
  insq = norm2(l, in, eo);
  if (insq == 0.) {
    for (int i = 0; i < ns; i++) {
      vzero(l, out[i], eo);
      ra[i]->final_rsq = 0; ra[i]->final_rel = 0; ra[i]->final_iter = 0; ra[i]->final_restart = 0;
    }
    return;
  }
  int imin = 0, irsq = 0, irel = 0;
  for (int i = 1; i < ns; i++) {
    if (shifts[i] < shifts[imin]) imin = i;
    if (ra[i]->rsqmin >= ra[irsq]->rsqmin) irsq = i;
    if (ra[i]->relmin > ra[irel]->relmin) irel = i;
  }
  double rsqstop = ra[irsq]->rsqmin * insq;
  double relstop = ra[irel]->relmin;
  double *r = vnew(l), *Mp = vnew(l), *pm[ns];
  vcopy(l, r, in, eo);
  for (int i = 0; i < ns; i++) {
    pm[i] = i == imin ? p : vnew(l);
    vzero(l, out[i], eo);
    vcopy(l, pm[i], r, eo);
    zo[i] = z[i] = 1;
    bo[i] = -1;
    a[i] = 0;
    b[i] = 0;
  }
  rsq = insq;
  relnorm2 = 1;
  int rsqCheckInterval = 100, nextRsqCheck = rsqCheckInterval;
  double trueRsqFac = 1;
  int relCheckInterval = 100, nextRelCheck = relCheckInterval;
  double trueRelFac = 1;
  double yAy = 0;
  double *y = vnew(l), *Ay = vnew(l), *ry = vnew(l);
  vcopy(l, ry, in, eo);
  while (1) {
    oldrsq = rsq;
    pkp = lin_norm2(lc, Mp, p);
    iteration++;
    if (pkp <= 0) break;
    b[imin] = rsq / pkp;
    vaxpy(l, out[imin], b[imin], p, out[imin], eo);
    vaxpy(l, r, -b[imin], Mp, r, eo);
    zn[imin] = 1;
    for (int i = 0; i < ns; i++) {
      if (i == imin) continue;
      zn[i] = z[i] * zo[i] * bo[imin];
      double c1 = b[imin] * a[imin] * (zo[i] - z[i]);
      c1 += zo[i] * bo[imin] * (1 + shifts[i] * b[imin]);
      if (c1 != 0.0) zn[i] /= c1; else zn[i] = 0;
      if (z[i] != 0.0) b[i] = b[imin] * zn[i] / z[i]; else zn[i] = b[i] = 0;
    }
    for (int i = 0; i < ns; i++)
      if (i != imin) vaxpy(l, out[i], b[i], pm[i], out[i], eo);
    rsq = norm2(l, r, eo);
    double rsqc = zn[irsq] * zn[irsq] * rsq;
    if (relstop > 0) relnorm2 = zn[irel] * zn[irel] * or_relnorm2(l, r, out[irel], eo);
    if (verbose > 1) printf("CGMS: iter %i rsq = %g\n", iteration, rsqc);
    if (iteration >= ia->restart || iteration >= ia->max_iter) break;
    if (rsqc * trueRsqFac < rsqstop) nextRsqCheck = iteration;
    if (relnorm2 * trueRelFac < relstop) nextRelCheck = iteration;
    if (nextRsqCheck <= iteration || (relstop > 0 && nextRelCheck <= iteration)) {
      if (yAy > 0) { /* A-orthogonalise x to y */
        double yr, yi;
        cdot(l, Ay, out[imin], eo, &yr, &yi);
        vcaxpy(l, out[imin], -yr / yAy, -yi / yAy, y, eo);
      }
      double xAx = lin_norm2(lc, Mp, out[imin]);
      iteration++;
      if (xAx <= 0) break;
      double xr, xi;
      cdot(l, out[imin], ry, eo, &xr, &xi);
      vcaxpy(l, y, xr / xAx, xi / xAx, out[imin], eo);
      yAy = lin_norm2(lc, Ay, y);
      iteration++;
      vaxpy(l, ry, -1.0, Ay, in, eo);
      vzero(l, out[imin], eo);
    }
    if (nextRsqCheck <= iteration) {
      double tr2, en;
      if (irsq == imin) {
        tr2 = norm2(l, ry, eo);
        vaxpy(l, Mp, 1.0, in, ry, eo);
        en = redot(l, y, Mp, eo);
      } else {
        double xAx = lin_norm2(lc, Mp, out[irsq]);
        iteration++;
        if (xAx <= 0) break;
        vaxpy(l, Mp, -1.0, in, Mp, eo);
        tr2 = norm2(l, Mp, eo);
        vaxpy(l, Mp, -1.0, in, Mp, eo);
        en = -redot(l, out[irsq], Mp, eo);
      }
      if (verbose) printf("CGMS: %i: rsq: %g  tr2: %g  enrm: %.16g\n", iteration, rsq, tr2, en);
      if (tr2 <= rsqstop) { rsq = tr2; break; }
      if (tr2 > 2 * rsq) break;
      trueRsqFac = tr2 / rsq;
      nextRsqCheck += rsqCheckInterval;
    }
    if (relstop > 0 && nextRelCheck <= iteration) {
      double xAx = lin_norm2(lc, Mp, out[irel]);
      iteration++;
      if (xAx <= 0) break;
      vaxpy(l, Mp, -1.0, in, Mp, eo);
      double rel2 = or_relnorm2(l, Mp, out[irel], eo);
      if (rel2 <= relstop) { relnorm2 = rel2; break; }
      trueRelFac = rel2 / relnorm2;
      nextRelCheck += relCheckInterval;
      relnorm2 = rel2;
    }
    a[imin] = rsq / oldrsq;
    vaxpy(l, p, a[imin], p, r, eo);
    for (int i = 0; i < ns; i++) {
      if (i == imin) continue;
      double c2 = z[i] * b[imin];
      if (c2 != 0.0) a[i] = a[imin] * zn[i] * b[i] / c2; else a[i] = 0;
    }
    for (int i = 0; i < ns; i++) {
      if (i == imin) continue;
      /* pm = zn r + a pm */
      long bb, ee;
      sub_range(l, eo, &bb, &ee);
      double zi = zn[i], ai = a[i];
      double *pmi = pm[i];
#pragma omp parallel for
      for (long k = 6 * bb; k < 6 * ee; k++) pmi[k] = ST(l, ai * pmi[k] + zi * r[k]);
    }
    for (int i = 0; i < ns; i++) { bo[i] = b[i]; zo[i] = z[i]; z[i] = zn[i]; }
  }
  vcopy(l, out[imin], y, eo);
  free(y); free(Ay); free(ry); free(r); free(Mp);
  for (int i = 0; i < ns; i++) if (i != imin) free(pm[i]);
  for (int i = 0; i < ns; i++) {
    double s = zn[i] * zn[i];
    ra[i]->final_rsq = s * trueRsqFac * rsq / insq;
    ra[i]->final_rel = s * trueRelFac * relnorm2;
    ra[i]->final_iter = iteration;
    ra[i]->final_restart = 0;
  }
}

/* ------------------------------------------------------------------ */
/* solve_multi + refine (lib/asqtad_invert_p.c:246-647)                */
/* ------------------------------------------------------------------ */
/* r[i] holds A*out[i] on the opposite subset (Aout) and the residual on
 * the input subset, as in the reference; Aout is kept separate here. */
static void refine(or_links_t *l, double **out, double *const *in, double **r,
                   double **Aout, double *r2, const double *m2, double *scale,
                   int no, int ineo, double **x, double **Ax, int nx,
                   const int *oi, const int *xi, int ni, double *tv) {
  int opeo = oppsub(ineo);
  double x2[nx + 1], Ax2[nx + 1];
  for (int i = 0; i < nx; i++) {
    or_dslash(l, 0, Ax[i], x[i], opeo, ineo);
    x2[i] = norm2(l, x[i], ineo);
    Ax2[i] = norm2(l, Ax[i], opeo);
  }
  for (int i = 0; i < ni; i++) {
    int ko = oi[i];
    int kx = nx > 0 ? xi[i] : 0;
    double sxr = 0, sxi = 0, d = 0, zr = 0, zi = 0, azr = 0, azi = 0;
    or_dslash(l, 0, Aout[ko], out[ko], opeo, ineo);
    if (nx > 0) {
      cdot(l, x[kx], out[ko], ineo, &zr, &zi);
      cdot(l, Ax[kx], Aout[ko], opeo, &azr, &azi);
      d = Ax2[kx] + m2[ko] * x2[kx];
    }
    if (d != 0) {
      double di = 1 / d, dim2 = di * m2[ko];
      double sr = dim2 * zr + di * azr, si = dim2 * zi + di * azi;
      vcaxpy(l, out[ko], -sr, -si, x[kx], ineo);
      vcaxpy(l, Aout[ko], -sr, -si, Ax[kx], opeo);
      cdot(l, x[kx], in[ko], ineo, &zr, &zi);
      sxr = di * zr; sxi = di * zi;
    }
    double o2, Ao2;
    cdot(l, out[ko], in[ko], ineo, &zr, &zi);
    o2 = norm2(l, out[ko], ineo);
    Ao2 = norm2(l, Aout[ko], opeo);
    d = Ao2 + m2[kx] * o2; /* the reference indexes m2 with kx here (:299) */
    if (d != 0) {
      double di = 1 / d;
      double sr = di * zr, si = di * zi;
      scale[2 * ko] = sr; scale[2 * ko + 1] = si;
      if (nx > 0) { /* z = sx / s */
        double den = sr * sr + si * si;
        double qr = (sxr * sr + sxi * si) / den, qi = (sxi * sr - sxr * si) / den;
        vcaxpy(l, out[ko], qr, qi, x[kx], ineo);
        vcaxpy(l, Aout[ko], qr, qi, Ax[kx], opeo);
      }
    } else {
      scale[2 * ko] = sxr; scale[2 * ko + 1] = sxi;
      if (nx > 0) {
        vcopy(l, out[ko], x[kx], ineo);
        vcopy(l, Aout[ko], Ax[kx], opeo);
      }
    }
  }
  /* r <- in - scale * A out */
  for (int i = 0; i < no; i++) {
    or_dslash(l, 0, tv, Aout[i], ineo, opeo);
    vaxpy(l, tv, -m2[i], out[i], tv, ineo);
    vcopy(l, r[i], in[i], ineo);
    vcaxpy(l, r[i], scale[2 * i], scale[2 * i + 1], tv, ineo);
    r2[i] = norm2(l, r[i], ineo);
  }
}

/* FP32 copy of an FP64 link object (QOP_FD_asqtad_create_L_from_L). */
static or_links_t *links_copy_single(const or_links_t *l) {
  or_links_t *f = calloc(1, sizeof *f);
  *f = *l;
  for (int k = 0; k < l->nlinks; k++) {
    f->dbl[k] = malloc(18 * l->V * sizeof(double));
    for (long i = 0; i < 18 * l->V; i++) f->dbl[k][i] = (double)(float)l->dbl[k][i];
  }
  f->single = 1;
  return f;
}

static void solve_multi(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t **ra,
                        const double *masses, double **out, double *const *in,
                        int nsolve, double *final_flop) {
  int ineo = ia->evenodd;
  int opeo = oppsub(ineo);
  double in2[nsolve], r2[nsolve], r2stop[nsolve], mass2[nsolve], m2s[nsolve];
  double scale[2 * nsolve];
  int oi[nsolve], xi[nsolve], nm = 0, iter = 0, nrestart = -1;
  int max_iter = ia->max_iter, max_restarts = ia->max_restarts;
  double *r[nsolve], *Aout[nsolve], *tv = vnew(l);
  for (int i = 0; i < nsolve; i++) {
    in2[i] = norm2(l, in[i], ineo);
    r2stop[i] = ra[i]->rsqmin * in2[i];
    mass2[i] = masses[i] * masses[i];
    oi[i] = i;
    int j = nm;
    while (j > 0 && mass2[i] < m2s[j - 1]) j--;
    if (j == nm || m2s[j] != mass2[i]) {
      for (int k = nm; k > j; k--) m2s[k] = m2s[k - 1];
      m2s[j] = mass2[i];
      nm++;
    }
    r[i] = vnew(l);
    Aout[i] = vnew(l);
  }
  for (int i = 0; i < nsolve; i++) {
    int j = 0;
    while (j < nm - 1 && m2s[j] != mass2[i]) j++;
    xi[i] = j;
  }
  or_resid_arg_t resarg0 = {0, 0, 0, 0, 0, 0};
  or_resid_arg_t *xresarg[nm];
  double *x[nm], *Ax[nm];
  double m20 = m2s[0];
  linop_ctx lc = {l, sqrt(m20), ineo};
  for (int i = 0; i < nm; i++) {
    m2s[i] = ST(l, m2s[i] - m20); /* QLA_Real shifts (:356) */
    x[i] = vnew(l);
    Ax[i] = vnew(l);
  }
  double *cgp = vnew(l);
  refine(l, out, in, r, Aout, r2, mass2, scale, nsolve, ineo, NULL, NULL, 0, oi, NULL, nsolve, tv);
  while (1) {
    for (int i = 0; i < nsolve; i++)
      ra[i]->final_rel = ra[i]->relmin > 0 ? or_relnorm2(l, r[i], out[i], ineo) : 1;
    int imax = 0, usingrel = 0;
    double r2max = 0, r2stopmax = 1, in2max = 1, rsqminfac = 1;
    for (int i = 0; i < nsolve; i++) {
      double r2i = r2[i], r2stopi = r2stop[i], in2i = in2[i], rsqminfaci = 1;
      int usingreli = 0;
      if (ra[i]->final_rel * r2stopi < r2i * ra[i]->relmin) {
        r2i = ra[i]->final_rel; r2stopi = ra[i]->relmin; in2i = 1;
        rsqminfaci = 1e-10; usingreli = 1;
      }
      double aa = r2i * r2stopmax, bb = r2max * r2stopi;
      if (aa > bb * 1.01 || (aa * 1.01 >= bb && mass2[i] < mass2[imax])) {
        imax = i; r2max = r2i; r2stopmax = r2stopi; in2max = in2i;
        rsqminfac = rsqminfaci; usingrel = usingreli;
      }
    }
    if (verbose)
      printf("SOLVE: its: %i worst: %i mass: %g rsq: %g/%g\n", iter, imax, masses[imax],
             r2[imax] / in2[imax], ra[imax]->rsqmin);
    if (r2max <= r2stopmax || nrestart >= max_restarts || iter >= max_iter) break;
    for (int i = 0; i < nm; i++) {
      xresarg[i] = i == xi[imax] ? ra[imax] : &resarg0;
      vzero(l, x[i], OR_EVENODD);
    }
    double rsqminold = ra[imax]->rsqmin;
    ra[imax]->rsqmin = r2stopmax / r2max;
    if (ra[imax]->rsqmin >= 1) ra[imax]->rsqmin = in2max * r2stopmax / r2max;
    ra[imax]->rsqmin *= rsqminfac;
    ia->max_iter = max_iter - iter;
    if (!l->single && r2max > ia->mixed_rsq * in2max) {
      /* FP32 inner multi-shift CG on an FP32 copy of the links, the residual
       * and the shifts, when r2 > mixed_rsq in2 (:510-538; the links copy is
       * QOP_FD_asqtad_create_L_from_L, asqtad_dslash_p.c:742-790) */
      or_links_t *lf = links_copy_single(l);
      if (ia->mixed_rsq > rsqminold) ra[imax]->rsqmin = ia->mixed_rsq * in2max / r2max;
      double m2sf[nm], *rf = vnew(l), *xf[nm], *cgpf = vnew(l);
      for (int i = 0; i < nm; i++) { m2sf[i] = (double)(float)m2s[i]; xf[i] = vnew(l); }
      for (long q = 0; q < 6 * l->V; q++) rf[q] = (double)(float)r[imax][q];
      linop_ctx lcf = {lf, sqrt(m20), ineo};
      invert_cgms(&lcf, ia, xresarg, m2sf, nm, xf, rf, cgpf);
      for (int i = 0; i < nm; i++) { vcopy(l, x[i], xf[i], ineo); free(xf[i]); }
      free(rf); free(cgpf);
      or_links_destroy(lf);
    } else {
      invert_cgms(&lc, ia, xresarg, m2s, nm, x, r[imax], cgp);
    }
    iter += ra[imax]->final_iter;
    nrestart++;
    ra[imax]->rsqmin = rsqminold;
    if (usingrel) {
      for (int i = 0; i < nsolve; i++) {
        int ko = oi[i], kx = xi[i];
        double s2 = scale[2 * ko] * scale[2 * ko] + scale[2 * ko + 1] * scale[2 * ko + 1];
        if (s2 == 0.0) vcopy(l, out[ko], x[kx], ineo);
        else vcaxpy(l, out[ko], scale[2 * ko] / s2, -scale[2 * ko + 1] / s2, x[kx], ineo);
      }
      refine(l, out, in, r, Aout, r2, mass2, scale, nsolve, ineo, NULL, NULL, 0, oi, NULL, nsolve, tv);
    } else {
      refine(l, out, in, r, Aout, r2, mass2, scale, nsolve, ineo, x, Ax, nm, oi, xi, nsolve, tv);
    }
  }
  ia->max_iter = max_iter;
  for (int i = 0; i < nsolve; i++) {
    double sr = masses[i] * scale[2 * i], si = masses[i] * scale[2 * i + 1];
    vzero(l, tv, ineo);
    vcaxpy(l, tv, sr, si, out[i], ineo);
    vcopy(l, out[i], tv, ineo);
    ra[i]->final_iter = iter;
    ra[i]->final_rsq = in2[i] > 0 ? r2[i] / in2[i] : 0;
    ra[i]->final_restart = nrestart > 0 ? nrestart : 0;
  }
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * l->nlinks);
  double nflopm = 0.5 * 10 * 3;
  if (final_flop) *final_flop = (nflop + nflopm * (nm - 1)) * iter * (double)l->V;
  for (int i = 0; i < nm; i++) { free(x[i]); free(Ax[i]); }
  for (int i = 0; i < nsolve; i++) { free(r[i]); free(Aout[i]); }
  free(tv); free(cgp);
  (void)opeo;
}

/* ------------------------------------------------------------------ */
/* QOP_asqtad_invert (asqtad_invert_p.c:649-851)                      */
/* ------------------------------------------------------------------ */
static void invert_eigcg(linop_ctx *lc, or_invert_arg_t *ia, or_resid_arg_t *ra, double *out,
                         const double *in, double *p, or_eigcg_t *eg);

static void invert_eo(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t *ra, double mass,
                      double *out, const double *in, double *final_flop, or_eigcg_t *eg);

void or_invert(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t *ra,
               double mass, double *out, const double *in, double *final_flop) {
  if (ia->evenodd != OR_EVENODD) {
    double *inp[1] = {(double *)in}, *outp[1] = {out};
    or_resid_arg_t *rap[1] = {ra};
    solve_multi(l, ia, rap, &mass, outp, inp, 1, final_flop);
    return;
  }
  invert_eo(l, ia, ra, mass, out, in, final_flop, NULL);
}

/* The EVENODD invert with QOP_asqtad.cgtype = 1: eigCG for the inner solves
 * (asqtad_invert_p.c:683-687, 767-769), state kept in eg across calls as
 * the reference keeps it in the links object. */
void or_invert_eigcg(or_links_t *l, or_eigcg_t *eg, or_invert_arg_t *ia, or_resid_arg_t *ra,
                     double mass, double *out, const double *in, double *final_flop) {
  invert_eo(l, ia, ra, mass, out, in, final_flop, eg);
}

static void invert_eo(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t *ra, double mass,
                      double *out, const double *in, double *final_flop, or_eigcg_t *eg) {
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * l->nlinks);
  int ineo = ia->evenodd, iter = 0, nrestart = -1;
  int max_iter_old = ia->max_iter, max_restarts_old = ia->max_restarts;
  int max_restarts = ia->max_restarts;
  if (max_restarts <= 0) max_restarts = 5;
  double insqe = 0, insqo = 0;
  if (ineo != OR_ODD) insqe = norm2(l, in, OR_EVEN);
  if (ineo != OR_EVEN) insqo = norm2(l, in, OR_ODD);
  double insq = insqe + insqo;
  int cgeo = insqe >= insqo ? OR_EVEN : OR_ODD;
  linop_ctx lc = {l, mass, cgeo};
  double *qin = vnew(l), *qout = vnew(l), *cgp = vnew(l), *cgr = vnew(l), *tmp2 = vnew(l), *gtmp = vnew(l);
  vcopy(l, tmp2, in, ineo);
  if (ineo != OR_EVENODD) vzero(l, tmp2, oppsub(ineo));
  /* project: qin = mass*tmp2 - D tmp2 on cgeo (:79-89) */
  or_dslash(l, mass, qin, tmp2, cgeo, oppsub(cgeo));
  {
    long b, e;
    sub_range(l, cgeo, &b, &e);
    for (long i = 6 * b; i < 6 * e; i++) qin[i] = ST(l, mass * tmp2[i] - qin[i]);
  }
  if (ineo == cgeo) vzero(l, qin, oppsub(cgeo));
  else vcopy(l, qin, in, oppsub(cgeo));
  double pinsq = norm2(l, qin, cgeo);
  double rsqfac = mass * mass * insq / pinsq;
  vzero(l, qout, OR_EVENODD);
  vcopy(l, qout, out, cgeo);
  double rsqstop = insq * ra->rsqmin, rsq = 0, relnorm2 = 1.;
  double rsqminold = ra->rsqmin, relminold = ra->relmin;
  ra->rsqmin = 0.99 * rsqminold * rsqfac;
  ra->relmin = 0.99 * relminold;
  ia->max_restarts = 0;
  do {
    if (eg) invert_eigcg(&lc, ia, ra, qout, qin, cgp, eg);
    else invert_cg(&lc, ia, ra, qout, qin, cgp);
    iter += ra->final_iter;
    /* reconstruct: qout_o = (qin_o - D qout_e)/mass on opp(cgeo) (:91-100) */
    int oe = oppsub(cgeo);
    vcopy(l, gtmp, qout, cgeo);
    or_dslash(l, mass, qout, gtmp, oe, cgeo);
    vaxpy(l, qout, -1.0, qout, qin, oe);
    {
      long b, e;
      sub_range(l, oe, &b, &e);
      double f = 1.0 / mass;
      for (long i = 6 * b; i < 6 * e; i++) qout[i] = ST(l, f * qout[i]);
    }
    /* full residual (:797-799) */
    or_dslash(l, mass, cgr, qout, ineo, OR_EVENODD);
    vaxpy(l, cgr, -1.0, in, cgr, ineo);
    rsq = norm2(l, cgr, ineo);
    ra->rsqmin *= 0.99;
    if (ra->relmin > 0) {
      relnorm2 = or_relnorm2(l, cgr, qout, ineo);
      ra->relmin = 0.99 * ra->final_rel / relnorm2 * relminold;
    }
    if (verbose) printf("ASQTAD_INVERT: iter %i rsq = %g\n", iter, rsq);
    ia->max_iter = max_iter_old - iter;
  } while (++nrestart < max_restarts && ia->max_iter > 0 && rsq > rsqstop &&
           relnorm2 > ra->relmin);
  vcopy(l, out, qout, ineo);
  ia->max_iter = max_iter_old;
  ia->max_restarts = max_restarts_old;
  ra->rsqmin = rsqminold;
  ra->relmin = relminold;
  ra->final_iter = iter;
  ra->final_rsq = insq == 0. ? 0 : rsq / insq;
  ra->final_rel = relnorm2;
  ra->final_restart = nrestart;
  if (final_flop) *final_flop = nflop * iter * (double)l->V;
  free(qin); free(qout); free(cgp); free(cgr); free(tmp2); free(gtmp);
}

/* ------------------------------------------------------------------ */
/* QOP_asqtad_invert_multi (asqtad_invert_p.c:878-1059)               */
/* ------------------------------------------------------------------ */
void or_invert_multi(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t *ra,
                     const double *masses, const int *nmass, double **out,
                     double *const *in, int nsrc, double *final_flop) {
  int nsolve = 0;
  for (int i = 0; i < nsrc; i++) nsolve += nmass[i];
  if (ia->evenodd != OR_EVENODD) {
    or_resid_arg_t *rap[nsolve];
    double *inp[nsolve];
    int k = 0;
    for (int i = 0; i < nsrc; i++)
      for (int j = 0; j < nmass[i]; j++) { rap[k] = &ra[k]; inp[k] = in[i]; k++; }
    solve_multi(l, ia, rap, masses, out, inp, nsolve, final_flop);
    return;
  }
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * l->nlinks), nflopm = 0.5 * 10 * 3;
  int cgeo = OR_EVEN; /* EVENODD is "not supported yet" and uses EVEN (:938-941) */
  double flop = 0;
  double *cgp = vnew(l);
  if (nsrc == 2 && nmass[0] == 1 && nmass[1] == 1 && masses[0] != masses[1]) {
    double shifts[2];
    int imin = 0;
    for (int i = 0; i < 2; i++) {
      shifts[i] = masses[i] * masses[i];
      if (shifts[i] < shifts[imin]) imin = i;
    }
    double gm = masses[imin];
    for (int i = 0; i < 2; i++) shifts[i] = ST(l, shifts[i] - gm * gm);
    linop_ctx lc = {l, gm, cgeo};
    double st = 1 / (shifts[1] - shifts[0]);
    double *x0 = vnew(l), *src = vnew(l);
    vaxpy(l, x0, -1.0, in[0], in[1], cgeo);
    {
      long b, e;
      sub_range(l, cgeo, &b, &e);
      for (long q = 6 * b; q < 6 * e; q++) x0[q] = ST(l, st * x0[q]);
    }
    vcopy(l, cgp, x0, cgeo);
    lin_plain(&lc, src, cgp);
    vaxpy(l, src, -1.0, src, in[imin], cgeo);
    double rsq1 = norm2(l, src, cgeo), rsqsave[2];
    or_resid_arg_t *rap[2] = {&ra[0], &ra[1]};
    for (int i = 0; i < 2; i++) {
      double rsq0 = norm2(l, in[i], cgeo);
      rsqsave[i] = ra[i].rsqmin;
      ra[i].rsqmin *= rsq0 / rsq1;
    }
    invert_cgms(&lc, ia, rap, shifts, 2, out, src, cgp);
    flop += (nflop + nflopm) * ra[imin].final_iter * (double)l->V;
    for (int i = 0; i < 2; i++) {
      ra[i].rsqmin = rsqsave[i];
      vaxpy(l, out[i], 1.0, x0, out[i], cgeo);
    }
    free(x0); free(src);
  } else {
    int k0 = 0;
    for (int i = 0; i < nsrc; i++) {
      int nm = nmass[i], jmin = 0;
      double shifts[nm];
      or_resid_arg_t *rap[nm];
      for (int j = 0; j < nm; j++) {
        shifts[j] = masses[k0 + j] * masses[k0 + j];
        if (shifts[j] < shifts[jmin]) jmin = j;
        rap[j] = &ra[k0 + j];
      }
      double gm = masses[k0 + jmin];
      for (int j = 0; j < nm; j++) shifts[j] = ST(l, shifts[j] - gm * gm);
      linop_ctx lc = {l, gm, cgeo};
      invert_cgms(&lc, ia, rap, shifts, nm, out + k0, in[i], cgp);
      flop += (nflop + nflopm * (nm - 1)) * ra[k0].final_iter * (double)l->V;
      k0 += nm;
    }
  }
  int k = 0;
  for (int i = 0; i < nsrc; i++)
    for (int j = 0; j < nmass[i]; j++, k++) {
      long b, e;
      sub_range(l, cgeo, &b, &e);
      for (long q = 6 * b; q < 6 * e; q++) out[k][q] = ST(l, out[k][q] * masses[k]);
    }
  if (final_flop) *final_flop = flop;
  free(cgp);
}

/* ------------------------------------------------------------------ */
/* eigCG (lib/invert_eigcg_p.c:1-1070, Stathopoulos and Orginos,       */
/* arXiv:0707.0131).  The small dense eigenproblems the reference     */
/* hands to LAPACK (dsyev, zheev: all eigenpairs, ascending) are       */
/* solved here by cyclic Jacobi; eigenvectors are defined up to sign   */
/* or phase, which the algorithm does not depend on.                   */
/* ------------------------------------------------------------------ */
struct or_eigcg_s {
  int numax, m, nev, nu, nv, nn, addvecs, cap, eo;
  double **u;
  double *l;
};

or_eigcg_t *or_eigcg_create(void) {
  or_eigcg_t *eg = calloc(1, sizeof(or_eigcg_t));
  eg->eo = -1;
  return eg;
}
void or_eigcg_destroy(or_eigcg_t *eg) {
  if (!eg) return;
  if (eg->u) {
    for (int i = 0; i < eg->cap; i++) free(eg->u[i]);
    free(eg->u);
    free(eg->l);
  }
  free(eg);
}
void or_eigcg_set(or_eigcg_t *eg, int numax, int m, int nev) {
  eg->numax = numax;
  eg->m = m;
  eg->nev = nev;
}
int or_eigcg_nu(const or_eigcg_t *eg, double *l, int maxn) {
  for (int i = 0; i < eg->nu && i < maxn; i++) l[i] = eg->l[i];
  return eg->nu;
}

/* real symmetric a (n x n row-major, destroyed): w ascending, z columns */
static void jacobi_sym(int n, double *a, double *w, double *z) {
  memset(z, 0, (size_t)n * n * sizeof(double));
  for (int i = 0; i < n; i++) z[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += a[i * n + j] * a[i * n + j];
        if (i != j) off += a[i * n + j] * a[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p * n + q];
        if (apq == 0) continue;
        double theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        double c = 1 / sqrt(t * t + 1), sn = t * c;
        for (int k = 0; k < n; k++) {
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - sn * akq;
          a[k * n + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - sn * aqk;
          a[q * n + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double zkp = z[k * n + p], zkq = z[k * n + q];
          z[k * n + p] = c * zkp - sn * zkq;
          z[k * n + q] = sn * zkp + c * zkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = a[i * n + i];
  /* sort ascending (selection, columns follow) */
  for (int i = 0; i < n; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (w[j] < w[k]) k = j;
    if (k != i) {
      double t = w[i]; w[i] = w[k]; w[k] = t;
      for (int r = 0; r < n; r++) {
        double x = z[r * n + i]; z[r * n + i] = z[r * n + k]; z[r * n + k] = x;
      }
    }
  }
}

/* complex Hermitian a (n x n row-major, re/im interleaved, destroyed) */
static void jacobi_herm(int n, double complex *a, double *w, double complex *z) {
  memset(z, 0, (size_t)n * n * sizeof(double complex));
  for (int i = 0; i < n; i++) z[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double v = creal(a[i * n + j] * conj(a[i * n + j]));
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double complex apq = a[p * n + q];
        double g = cabs(apq);
        if (g == 0) continue;
        double complex ph = apq / g; /* a_pq = g e^{i phi} */
        double theta = (creal(a[q * n + q]) - creal(a[p * n + p])) / (2 * g);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        double c = 1 / sqrt(t * t + 1), sn = t * c;
        /* J: columns p' = c e_p - sn conj(ph) e_q, q' = sn ph e_p + c e_q,
         * A <- J^+ A J zeroes a_pq */
        for (int k = 0; k < n; k++) {
          double complex akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - sn * conj(ph) * akq;
          a[k * n + q] = sn * ph * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double complex apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - sn * ph * aqk;
          a[q * n + k] = sn * conj(ph) * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double complex zkp = z[k * n + p], zkq = z[k * n + q];
          z[k * n + p] = c * zkp - sn * conj(ph) * zkq;
          z[k * n + q] = sn * ph * zkp + c * zkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = creal(a[i * n + i]);
  for (int i = 0; i < n; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (w[j] < w[k]) k = j;
    if (k != i) {
      double t = w[i]; w[i] = w[k]; w[k] = t;
      for (int r = 0; r < n; r++) {
        double complex x = z[r * n + i]; z[r * n + i] = z[r * n + k]; z[r * n + k] = x;
      }
    }
  }
}

typedef struct {
  int m, nev;
  double *t, *ta, *tb, *td, *y, *h; /* m x m, m x m, m x m, m, m x 2nev, 2nev x 2nev (row-major) */
} eg_temps;

/* v'_j = sum_i R(i, j) v_i for j < n2 (R n1 x n2, row stride ld, complex or real) */
static void rotate_vecs(linop_ctx *lc, double **v, int n1, int n2, const double *Rr,
                        const double complex *Rz, int ld) {
  or_links_t *l = lc->l;
  double **tv = malloc(n2 * sizeof(double *));
  for (int j = 0; j < n2; j++) {
    tv[j] = vnew(l);
    for (int i = 0; i < n1; i++) {
      double cr = Rr ? Rr[i * ld + j] : creal(Rz[i * ld + j]);
      double ci = Rr ? 0.0 : cimag(Rz[i * ld + j]);
      vcaxpy(l, tv[j], cr, ci, v[i], lc->eo);
    }
  }
  for (int j = 0; j < n2; j++) {
    vcopy(l, v[j], tv[j], lc->eo);
    free(tv[j]);
  }
  free(tv);
}

static void normalize1(linop_ctx *lc, double *v) {
  const double r = 1.0 / sqrt(norm2(lc->l, v, lc->eo));
  long b, e;
  sub_range(lc->l, lc->eo, &b, &e);
  for (long i = 6 * b; i < 6 * e; i++) v[i] *= r;
}

/* invert_eigcg_p.c:201-213 */
static void orthonormalize_vecs(linop_ctx *lc, double **v, int i0, int n) {
  for (int i = i0; i < i0 + n; i++) {
    for (int j = 0; j < i; j++) {
      double zr, zi;
      cdot(lc->l, v[j], v[i], lc->eo, &zr, &zi);
      vcaxpy(lc->l, v[i], -zr, -zi, v[j], lc->eo);
    }
    normalize1(lc, v[i]);
  }
}

/* invert_eigcg_p.c:215-326 (the evrsq = NULL branch) */
static void rayleigh_ritz(linop_ctx *lc, double **u, double *lv, int nu, int nv, double *Mp) {
  int n = nu + nv;
  if (n <= 0) return;
  orthonormalize_vecs(lc, u, nu, nv);
  double complex *mm = calloc((size_t)n * n, sizeof(double complex));
  double complex *tt = malloc((size_t)n * n * sizeof(double complex));
  double *vv = malloc(n * sizeof(double));
  for (int i = 0; i < nu; i++) mm[i * n + i] = lv[i];
  for (int j = nu; j < n; j++) {
    lin_plain(lc, Mp, u[j]);
    for (int i = 0; i < j; i++) {
      double zr, zi;
      cdot(lc->l, u[i], Mp, lc->eo, &zr, &zi);
      mm[i * n + j] = zr + I * zi;
      mm[j * n + i] = zr - I * zi;
    }
    double zr, zi;
    cdot(lc->l, u[j], Mp, lc->eo, &zr, &zi);
    mm[j * n + j] = zr + I * zi;
  }
  jacobi_herm(n, mm, vv, tt);
  rotate_vecs(lc, u, n, n, NULL, tt, n);
  for (int i = 0; i < n; i++) lv[i] = vv[i];
  if (n == 1) lv[1] = 0;
  free(mm); free(tt); free(vv);
}

/* invert_eigcg_p.c:328-341 */
static void initcg(linop_ctx *lc, double *x, double **u, const double *lv, int n, const double *b) {
  if (n <= 0) return;
  vzero(lc->l, x, lc->eo);
  for (int i = 0; i < n; i++) {
    double zr, zi;
    cdot(lc->l, u[i], b, lc->eo, &zr, &zi);
    vcaxpy(lc->l, x, zr / lv[i], zi / lv[i], u[i], lc->eo);
  }
}

/* t(i, j) = Re <v_i, A v_j> (invert_eigcg_p.c:363-394) */
static void gett(linop_ctx *lc, double *t, int ldt, double **v, int n, double *Mp) {
  for (int j = 0; j < n; j++) {
    lin_plain(lc, Mp, v[j]);
    for (int i = 0; i < n; i++) t[i * ldt + j] = redot(lc->l, v[i], Mp, lc->eo);
  }
}

/* td[0:nv], tb[0:nv,0:nv] = eigenpairs of t[0:nv,0:nv] (:488-501) */
static void diag_t(eg_temps *et, int nv) {
  int m = et->m;
  double *a = malloc((size_t)nv * nv * sizeof(double)), *z = malloc((size_t)nv * nv * sizeof(double));
  for (int i = 0; i < nv; i++)
    for (int j = 0; j < nv; j++) a[i * nv + j] = et->t[i * m + j];
  jacobi_sym(nv, a, et->td, z);
  for (int i = 0; i < nv; i++)
    for (int j = 0; j < nv; j++) et->tb[i * m + j] = z[i * nv + j];
  free(a); free(z);
}

/* invert_eigcg_p.c:503-550 */
static void set_y(eg_temps *et, int noff, int nev, int nv, int n) {
  int m = et->m, ly = 2 * et->nev;
  double *y = et->y;
  for (int i = 0; i < nv; i++)
    for (int j = 0; j < nev; j++) y[i * ly + noff + j] = et->tb[i * m + j];
  for (int i = nv; i < n; i++)
    for (int j = 0; j < nev; j++) y[i * ly + noff + j] = 0;
  if (!noff) return;
  /* dorthocols: columns noff.. against columns 0..noff-1 and each other */
  for (int c = noff; c < noff + nev; c++) {
    for (int k = 0; k < c; k++) {
      double d = 0;
      for (int i = 0; i < n; i++) d += y[i * ly + k] * y[i * ly + c];
      for (int i = 0; i < n; i++) y[i * ly + c] -= d * y[i * ly + k];
    }
    double s = 0;
    for (int i = 0; i < n; i++) s += y[i * ly + c] * y[i * ly + c];
    s = 1 / sqrt(s);
    for (int i = 0; i < n; i++) y[i * ly + c] *= s;
  }
  int k = noff + nev;
  /* tb1 = t y4 (n x k); h0 = y4^T tb1; eig -> td, ta; tb1 = y4 ta */
  double *tb1 = calloc((size_t)n * k, sizeof(double));
  for (int i = 0; i < n; i++)
    for (int j = 0; j < k; j++) {
      double s2 = 0;
      for (int q = 0; q < n; q++) s2 += et->t[i * m + q] * y[q * ly + j];
      tb1[i * k + j] = s2;
    }
  double *h0 = calloc((size_t)k * k, sizeof(double)), *ta = malloc((size_t)k * k * sizeof(double));
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) {
      double s2 = 0;
      for (int q = 0; q < n; q++) s2 += y[q * ly + i] * tb1[q * k + j];
      h0[i * k + j] = s2;
    }
  jacobi_sym(k, h0, et->td, ta);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < k; j++) {
      double s2 = 0;
      for (int q = 0; q < k; q++) s2 += y[i * ly + q] * ta[q * k + j];
      et->tb[i * m + j] = s2;
    }
  free(tb1); free(h0); free(ta);
}

/* invert_eigcg_p.c:552-559: v[0:n] = v[0:nv] tb[0:nv, 0:n] */
static void rot_t(linop_ctx *lc, eg_temps *et, int n, double **v, int nv) {
  if (n > nv) n = nv;
  rotate_vecs(lc, v, nv, n, et->tb, NULL, et->m);
}

static void setv(linop_ctx *lc, double **v, int *nv, const double *r, double rsq) {
  vzero(lc->l, v[*nv], lc->eo);
  vaxpy(lc->l, v[*nv], 1 / sqrt(rsq), r, v[*nv], lc->eo);
  (*nv)++;
}

/* invert_eigcg_p.c:561-616 */
static void reset_t(linop_ctx *lc, eg_temps *et, int nev, double **v, int *nvp, double a0, double a,
                    double b, const double *r, double rsq, double *Mp) {
  int m = et->m, nv = 2 * nev;
  setv(lc, v, &nv, r, rsq);
  memset(et->t, 0, (size_t)m * m * sizeof(double));
  lin_plain(lc, Mp, v[nv - 1]);
  for (int i = 0; i < nv - 1; i++) {
    double rr = redot(lc->l, v[i], Mp, lc->eo);
    et->t[i * m + nv - 1] = rr;
    et->t[(nv - 1) * m + i] = rr;
    et->t[i * m + i] = et->td[i];
  }
  et->t[(nv - 1) * m + nv - 1] = 1. / a + b / a0;
  *nvp = nv;
}

/* QOP_invert_eigcg (invert_eigcg_p.c:618-1068) */
static void invert_eigcg(linop_ctx *lc, or_invert_arg_t *ia, or_resid_arg_t *ra, double *out,
                         const double *in, double *p, or_eigcg_t *eg) {
  or_links_t *l = lc->l;
  const int eo = lc->eo;
  double a = 0, b = 0, a0 = 0, rsq, relnorm2, oldrsq, pkp, insq, rsqstop;
  int iteration = 0, total = 0, nrestart = -1;
  int max_iterations = ia->max_iter, max_restarts = ia->max_restarts;
  if (max_restarts < 0) max_restarts = 5;
  const int nev = eg->nev, m = eg->m, numax = eg->numax;
  if (!eg->u) {
    eg->cap = numax;
    eg->u = malloc(numax * sizeof(double *));
    eg->l = calloc(numax + 1, sizeof(double));
    for (int i = 0; i < numax; i++) eg->u[i] = vnew(l);
    eg->nu = eg->nv = eg->nn = 0;
    eg->addvecs = 1;
  }
  if (eg->eo != eo) {
    /* the vectors live on the parity of the solves that made them; a solve
     * on the other parity starts a new space (the reference would
     * orthonormalise vectors that are zero there) */
    eg->nu = eg->nv = eg->nn = 0;
    eg->addvecs = 1;
    eg->eo = eo;
  }
  double **u = eg->u, *lv = eg->l;
  int nu = eg->nu, nv = eg->nv, nn = eg->nn, addvecs = eg->addvecs;
  /* the reference writes the Lanczos vectors to u[nu + nn ..] without a
   * bound; stop collecting when they would not fit */
  if (addvecs && nu + nn + m > eg->cap) addvecs = 0;
  eg_temps et = {m, nev, calloc((size_t)m * m, 8), calloc((size_t)m * m, 8), calloc((size_t)m * m, 8),
                 calloc(m + 1, 8), calloc((size_t)m * 2 * nev, 8), calloc((size_t)4 * nev * nev, 8)};
  double *r = vnew(l), *Mp = vnew(l), *Mq = vnew(l);
  double **v = NULL;
  insq = norm2(l, in, eo);
  rsqstop = ra->rsqmin * insq;
  rsq = 0;
  relnorm2 = 1.;
  ra->final_rsq = 0; ra->final_rel = 0; ra->final_iter = 0; ra->final_restart = 0;
  if (insq == 0.) {
    vzero(l, out, eo);
    goto done;
  }
  do {
    vcopy(l, p, out, eo);
    lin_plain(lc, Mp, p);
    iteration = 1;
    total++;
    vaxpy(l, r, -1.0, Mp, in, eo);
    rsq = norm2(l, r, eo);
    if (((rsqstop <= 0 || rsq < rsqstop) && (ra->relmin <= 0 || relnorm2 < ra->relmin)) ||
        total >= max_iterations)
      break;
    if (addvecs && nn >= m) {
      int nr = numax - 2 * m - 1;
      if (nr >= nu) nr = nu - 2;
      if (nr < 0) nr = 0;
      double oldl = nu == 0 ? 9999 : lv[nr];
      rayleigh_ritz(lc, u, lv, nu, nn, Mq);
      nu += nn;
      nn = 0;
      double delta = fabs(2 * (oldl - lv[nr]) / (oldl + lv[nr]));
      if (delta < 1e-4) addvecs = 0;
      else if (nu > numax - 2 * m) nu = numax - 2 * m;
    }
    if (nu > 0) {
      initcg(lc, p, u, lv, nu, r);
      vaxpy(l, out, 1.0, p, out, eo);
      vcopy(l, p, out, eo);
      lin_plain(lc, Mp, p);
      iteration = 1;
      total++;
      vaxpy(l, r, -1.0, Mp, in, eo);
      rsq = norm2(l, r, eo);
      if (ra->relmin > 0) relnorm2 = or_relnorm2(l, r, out, eo);
      if (((rsqstop <= 0 || rsq < rsqstop) && (ra->relmin <= 0 || relnorm2 < ra->relmin)) ||
          total >= max_iterations)
        break;
    }
    if (addvecs) {
      v = u + nu + nn;
      memset(et.t, 0, (size_t)m * m * sizeof(double));
    }
    vcopy(l, p, r, eo);
    while (1) {
      lin_plain(lc, Mp, p);
      iteration++;
      total++;
      pkp = redot(l, p, Mp, eo);
      a0 = a;
      a = rsq / pkp;
      if (addvecs) {
        if (nv == 0) {
          et.t[0] = 1. / a;
          setv(lc, v, &nv, r, rsq);
        } else if (nv < m) {
          double td = 1. / a + b / a0, ts = -sqrt(b) / a0;
          et.t[nv * m + nv] = td;
          et.t[nv * m + nv - 1] = ts;
          et.t[(nv - 1) * m + nv] = ts;
          setv(lc, v, &nv, r, rsq);
        } else {
          diag_t(&et, nv);
          set_y(&et, 0, nev, nv, nv);
          diag_t(&et, nv - 1);
          set_y(&et, nev, nev, nv - 1, nv);
          rot_t(lc, &et, 2 * nev, v, nv);
          reset_t(lc, &et, nev, v, &nv, a0, a, b, r, rsq, Mq);
        }
      }
      vaxpy(l, out, a, p, out, eo);
      vaxpy(l, r, -a, Mp, r, eo);
      oldrsq = rsq;
      rsq = norm2(l, r, eo);
      if (ra->relmin > 0) relnorm2 = or_relnorm2(l, r, out, eo);
      if (((rsqstop <= 0 || rsq < rsqstop) && (ra->relmin <= 0 || relnorm2 < ra->relmin)) ||
          total >= max_iterations)
        break;
      b = rsq / oldrsq;
      vaxpy(l, p, b, p, r, eo);
    }
    if (addvecs) {
      if (nv > nev) {
        orthonormalize_vecs(lc, v, 0, nv);
        gett(lc, et.t, m, v, nv, Mq);
        diag_t(&et, nv);
        set_y(&et, 0, nev, nv, nv);
        rot_t(lc, &et, nev, v, nv);
        nn += nev;
      }
      nv = 0;
    }
    nrestart++;
  } while (total < max_iterations && nrestart < max_restarts);
  if (addvecs) {
    rayleigh_ritz(lc, u, lv, 0, nu + nn + nv, Mq);
    nu += nn + nv;
    nn = 0;
    nv = 0;
  }
  eg->nu = nu;
  eg->nv = nv;
  eg->nn = nn;
  eg->addvecs = addvecs;
  ra->final_rsq = rsq / insq;
  ra->final_rel = relnorm2;
  ra->final_iter = total;
  ra->final_restart = nrestart;
done:
  free(r); free(Mp); free(Mq);
  free(et.t); free(et.ta); free(et.tb); free(et.td); free(et.y); free(et.h);
}
