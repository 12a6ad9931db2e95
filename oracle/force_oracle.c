/*
 * CPU oracle (TEST INFRASTRUCTURE ONLY): the asqtad fermion force of the
 * reference, "fnmat" form, restated in plain C on full-lattice matrix fields.
 *
 *   QOP_asqtad_force_multi_fnmat_qdp   lib/asqtad_force_fnmat_p.c:515-546
 *   QOP_asqtad_deriv_multi_fnmat_qdp   lib/asqtad_force_fnmat_p.c:474-513
 *   QOP_get_mid                        lib/asqtad_force_fnmat_p.c:4-44
 *   QOP_asqtad_deriv (Naik part)       lib/asqtad_force_fnmat_p.c:358-472
 *   QOP_fat_deriv                      lib/asqtad_force_fnmat_p.c:201-356
 *   staple / side_force                lib/asqtad_force_fnmat_p.c:93-199
 *
 * Every QDP operation of the reference becomes one call of a field primitive
 * below (fields are double[V][18], raw site order, shifts by coordinates):
 * QDP_M_eq_sM(d, s, mu, forward) -> d(x) = s(x + mu), backward -> s(x - mu).
 *
 * Parity pin (tests/test_oracle_force.py): the force is the derivative of
 * S(U) = sum_i eps_i |D_oe(U) X_i|^2 with respect to U_mu(z) -> exp(i t T)
 * U_mu(z) (through the fat7 + Lepage + Naik smearing), checked by central
 * finite differences on random sites, directions and generators, for the
 * MILC coefficients and for single-term coefficient sets.  The trace part of
 * the antihermitian projection (QDP_M_eq_antiherm_M) is taken as traceless;
 * the finite differences along traceless generators do not see it.  The
 * Lepage fix-up (:317-339) adds acc U - U acc(x+mu) with acc Hermitian, so
 * U deriv^+ gains a Hermitian term and the projected force does not depend
 * on it: it is restated for the derivative but not pinned.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "qop_oracle.h"

#define RE(m, i, j) (m)[6 * (i) + 2 * (j)]
#define IM(m, i, j) (m)[6 * (i) + 2 * (j) + 1]

typedef struct {
  int L[4];
  long V;
  long *fw[4], *bw[4];  // neighbour index tables x +- mu
} geom_t;

static void geom_init(geom_t *g, const int L[4]) {
  memcpy(g->L, L, sizeof g->L);
  g->V = or_volume(L);
  for (int mu = 0; mu < 4; mu++) {
    g->fw[mu] = malloc(g->V * sizeof(long));
    g->bw[mu] = malloc(g->V * sizeof(long));
  }
  for (long i = 0; i < g->V; i++) {
    int x[4];
    or_raw_coords(L, i, x);
    for (int mu = 0; mu < 4; mu++) {
      int y[4] = {x[0], x[1], x[2], x[3]};
      y[mu] = (x[mu] + 1) % L[mu];
      g->fw[mu][i] = or_raw_index(L, y);
      y[mu] = (x[mu] - 1 + L[mu]) % L[mu];
      g->bw[mu][i] = or_raw_index(L, y);
    }
  }
}
static void geom_free(geom_t *g) {
  for (int mu = 0; mu < 4; mu++) {
    free(g->fw[mu]);
    free(g->bw[mu]);
  }
}

static double *fnew(const geom_t *g) { return calloc(g->V * 18, sizeof(double)); }
static void fzero(const geom_t *g, double *a) { memset(a, 0, g->V * 18 * sizeof(double)); }
static void fcopy(const geom_t *g, double *d, const double *s) { memcpy(d, s, g->V * 18 * sizeof(double)); }

/* d(x) = s(x + dir*mu), dir = +1 forward, -1 backward */
static void fshift(const geom_t *g, double *d, const double *s, int mu, int dir) {
  const long *t = dir > 0 ? g->fw[mu] : g->bw[mu];
#pragma omp parallel for
  for (long i = 0; i < g->V; i++) memcpy(d + 18 * i, s + 18 * t[i], 18 * sizeof(double));
}

/* element (i,j) of op(a): adjoint when ta */
static inline double complex el(const double *a, int i, int j, int ta) {
  return ta ? RE(a, j, i) - I * IM(a, j, i) : RE(a, i, j) + I * IM(a, i, j);
}
/* c (+)= s * op(a) op(b) at one site */
static inline void smul(double *c, const double *a, int ta, const double *b, int tb, double s,
                        int acc) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double complex z = 0;
      for (int k = 0; k < 3; k++) z += el(a, i, k, ta) * el(b, k, j, tb);
      z *= s;
      RE(t, i, j) = creal(z);
      IM(t, i, j) = cimag(z);
    }
  if (acc)
    for (int e = 0; e < 18; e++) c[e] += t[e];
  else
    memcpy(c, t, sizeof t);
}
/* C = op(A) op(B) */
static void fmul(const geom_t *g, double *c, const double *a, int ta, const double *b, int tb) {
#pragma omp parallel for
  for (long i = 0; i < g->V; i++) smul(c + 18 * i, a + 18 * i, ta, b + 18 * i, tb, 1.0, 0);
}
/* C += s op(A) op(B) */
static void fmul_acc(const geom_t *g, double *c, const double *a, int ta, const double *b, int tb,
                     double s) {
#pragma omp parallel for
  for (long i = 0; i < g->V; i++) smul(c + 18 * i, a + 18 * i, ta, b + 18 * i, tb, s, 1);
}
/* C += s op(A) */
static void faxpy(const geom_t *g, double *c, double s, const double *a, int ta) {
#pragma omp parallel for
  for (long i = 0; i < g->V; i++)
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) {
        double complex z = el(a + 18 * i, r, k, ta);
        RE(c + 18 * i, r, k) += s * creal(z);
        IM(c + 18 * i, r, k) += s * cimag(z);
      }
}
/* C = s A */
static void fscale(const geom_t *g, double *c, double s, const double *a) {
#pragma omp parallel for
  for (long i = 0; i < 18 * g->V; i++) c[i] = s * a[i];
}

/* QOP_get_mid (:4-44): mid_s(x) = scale sum_i eps_i x_i(x) x_i(x + shift_s)^+;
 * shift_s = s-th direction times `dist` (1: fat, 3: Naik) */
static void get_mid(const geom_t *g, double *mid[4], int dist, const double eps[], double scale,
                    double *const x[], int nterms) {
  for (int s = 0; s < 4; s++) {
    fzero(g, mid[s]);
#pragma omp parallel for
    for (long i = 0; i < g->V; i++) {
      long j = i;
      for (int d = 0; d < dist; d++) j = g->fw[s][j];
      for (int t = 0; t < nterms; t++) {
        const double *a = x[t] + 6 * i, *b = x[t] + 6 * j;
        const double c = scale * eps[t];
        for (int r = 0; r < 3; r++)
          for (int k = 0; k < 3; k++) {
            const double complex z = (a[2 * r] + I * a[2 * r + 1]) * (b[2 * k] - I * b[2 * k + 1]);
            RE(mid[s] + 18 * i, r, k) += c * creal(z);
            IM(mid[s] + 18 * i, r, k) += c * cimag(z);
          }
      }
    }
  }
}

typedef struct {
  const geom_t *g;
  double *t[12];  // scratch fields
} ws_t;

/* staple (:93-127): out += link in(x+nu) link(x+mu)^+ + [link^+ in link(x+mu)](x-nu)
 * with link = link0 (direction nu), in = in0 (direction mu) */
static void staple(ws_t *w, double *out, const double *in0, const double *link0, int mu, int nu) {
  const geom_t *g = w->g;
  double *linkmu = w->t[0], *innu = w->t[1], *linkin = w->t[2], *back = w->t[3];
  double *backnu = w->t[4], *linkinnu = w->t[5];
  fshift(g, linkmu, link0, mu, +1);
  fshift(g, innu, in0, nu, +1);
  fmul(g, linkin, link0, 1, in0, 0);
  fmul(g, back, linkin, 0, linkmu, 0);
  fshift(g, backnu, back, nu, -1);
  fmul(g, linkinnu, link0, 0, innu, 0);
  fmul_acc(g, out, linkinnu, 0, linkmu, 1, 1.0);
  faxpy(g, out, 1.0, backnu, 0);
}

/* side_force (:129-199) */
static void side_force(ws_t *w, double *force, const double *bot, const double *side,
                       const double *top, int mu, int nu, double *stpl) {
  const geom_t *g = w->g;
  double *sidemu = w->t[0], *topnu = w->t[1], *botnu = w->t[2], *sidebot = w->t[3];
  double *sidetop = w->t[4], *tmp = w->t[5], *fbmu = w->t[6], *fmbmu = w->t[7];
  double *stm = w->t[8], *botnusidemu = w->t[9], *botsidemu = w->t[10];
  fshift(g, sidemu, side, mu, +1);
  fshift(g, topnu, top, nu, +1);
  fshift(g, botnu, bot, nu, +1);
  fmul(g, sidebot, side, 1, bot, 0);
  fmul(g, sidetop, side, 1, top, 0);
  fmul(g, tmp, topnu, 1, sidebot, 0);  // topnusidebot
  fshift(g, fbmu, tmp, mu, -1);
  fmul(g, tmp, botnu, 1, sidetop, 0);  // botnusidetop
  fshift(g, fmbmu, tmp, mu, -1);
  fmul(g, tmp, sidebot, 0, sidemu, 0);  // sidebotsidemu
  fshift(g, stm, tmp, nu, -1);
  fmul(g, botnusidemu, botnu, 0, sidemu, 1);
  fmul_acc(g, stpl, side, 0, botnusidemu, 0, 1.0);
  fmul_acc(g, force, top, 0, botnusidemu, 1, 1.0);
  fmul(g, botsidemu, bot, 0, sidemu, 0);
  fmul_acc(g, force, botsidemu, 0, topnu, 1, 1.0);
  faxpy(g, force, 1.0, fbmu, 1);
  faxpy(g, force, 1.0, fmbmu, 1);
  faxpy(g, stpl, 1.0, stm, 0);
}

/* QOP_fat_deriv (:201-356) */
static void fat_deriv(ws_t *w, double *const gauge[4], double *deriv[4],
                      const or_asqtad_coeffs_t *coef, double *const mid[4]) {
  const geom_t *g = w->g;
  double coef1 = coef->one_link, coef3 = coef->three_staple, coef5 = coef->five_staple;
  double coef7 = coef->seven_staple, coefL = coef->lepage;
  coef1 -= 6 * coefL;
  const int have5 = coef5 != 0 || coef7 != 0 || coefL != 0;
  const int have3 = coef3 != 0 || have5;
  double *stpl3[4] = {0}, *mid5[4] = {0}, *stpl5 = fnew(g), *mid3 = fnew(g);
  for (int mu = 0; mu < 4; mu++) {
    stpl3[mu] = fnew(g);
    mid5[mu] = fnew(g);
  }
  for (int sig = 0; sig < 4; sig++) {
    if (have5) {
      for (int mu = 0; mu < 4; mu++) {
        if (mu == sig) continue;
        fzero(g, stpl3[mu]);
        staple(w, stpl3[mu], gauge[sig], gauge[mu], sig, mu);
        fzero(g, mid5[mu]);
      }
      for (int rho = 0; rho < 4; rho++) {
        if (rho == sig) continue;
        fzero(g, stpl5);
        if (coef7 != 0) {
          for (int mu = 0; mu < 4; mu++) {
            if (mu == sig || mu == rho) continue;
            for (int nu = 0; nu < 4; nu++) {
              if (nu == sig || nu == rho || nu == mu) continue;
              staple(w, stpl5, stpl3[mu], gauge[nu], sig, nu);
            }
          }
          fscale(g, stpl5, coef7, stpl5);
        }
        fzero(g, mid3);
        if (coefL != 0) faxpy(g, stpl5, coefL, stpl3[rho], 0);
        if (coefL != 0 || coef7 != 0) side_force(w, deriv[rho], mid[sig], gauge[rho], stpl5, sig, rho, mid3);
        if (coefL != 0) faxpy(g, mid5[rho], coefL, mid3, 0);
        fscale(g, mid3, coef7, mid3);
        faxpy(g, mid3, coef5, mid[sig], 0);
        for (int mu = 0; mu < 4; mu++) {
          if (mu == sig || mu == rho) continue;
          for (int nu = 0; nu < 4; nu++) {
            if (nu == sig || nu == rho || nu == mu) continue;
            side_force(w, deriv[mu], mid3, gauge[mu], stpl3[nu], sig, mu, mid5[nu]);
          }
        }
      }
    }
    if (have3) {
      fzero(g, mid3);
      for (int mu = 0; mu < 4; mu++) {
        if (mu == sig) continue;
        if (have5) {
          fscale(g, stpl5, coef3, mid[sig]);
          faxpy(g, stpl5, 1.0, mid5[mu], 0);
        } else {
          fscale(g, stpl5, coef3, mid[sig]);
        }
        side_force(w, deriv[mu], stpl5, gauge[mu], gauge[sig], sig, mu, mid3);
      }
      faxpy(g, deriv[sig], 1.0, mid3, 0);
    }
    if (coef1 != 0) faxpy(g, deriv[sig], coef1, mid[sig], 0);
  }
  if (coefL != 0) {
    // fix up the Lepage term (:317-339)
    const double fixL = -coefL;
    double *acc = w->t[11], *tmp = w->t[10], *sh = w->t[9];
    for (int mu = 0; mu < 4; mu++) {
      fzero(g, acc);
      for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        fmul(g, tmp, mid[nu], 1, gauge[nu], 0);
        fshift(g, sh, tmp, nu, -1);
        fmul(g, stpl5, mid[nu], 0, gauge[nu], 1);
        faxpy(g, stpl5, -1.0, sh, 0);
        faxpy(g, acc, 1.0, stpl5, 0);
        faxpy(g, acc, 1.0, stpl5, 1);
      }
      fshift(g, sh, acc, mu, +1);
      fmul(g, stpl5, acc, 0, gauge[mu], 0);
      fmul_acc(g, stpl5, gauge[mu], 0, sh, 0, -1.0);
      faxpy(g, deriv[mu], fixL, stpl5, 0);
    }
  }
  for (int mu = 0; mu < 4; mu++) {
    free(stpl3[mu]);
    free(mid5[mu]);
  }
  free(stpl5);
  free(mid3);
}

/* Naik part of QOP_asqtad_deriv (:395-442) */
static void naik_deriv(ws_t *w, double *const gauge[4], double *deriv[4], double coefN,
                       double *const mid[4]) {
  const geom_t *g = w->g;
  double *Umu = w->t[0], *Umid = w->t[1], *Umidbmu = w->t[2], *UmuU = w->t[3];
  double *UmuUs = w->t[4], *f3b = w->t[5], *f3 = w->t[6], *f = w->t[7];
  for (int mu = 0; mu < 4; mu++) {
    const double *U = gauge[mu], *m = mid[mu];
    fshift(g, Umu, U, mu, +1);
    fmul(g, Umid, U, 1, m, 0);
    fshift(g, Umidbmu, Umid, mu, -1);
    fmul(g, UmuU, Umu, 1, U, 1);
    fshift(g, UmuUs, UmuU, mu, +1);
    fmul(g, f3b, U, 1, Umidbmu, 0);
    fshift(g, f3, f3b, mu, -1);
    fmul(g, f, m, 0, UmuUs, 0);
    fmul_acc(g, f, Umidbmu, 0, Umu, 1, 1.0);
    faxpy(g, f, 1.0, f3, 0);
    faxpy(g, deriv[mu], coefN, f, 0);
  }
}

/* traceless antihermitian part: (a - a^+)/2 - tr((a - a^+)/2)/3 */
static void antiherm(double *r, const double *a) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      RE(t, i, j) = 0.5 * (RE(a, i, j) - RE(a, j, i));
      IM(t, i, j) = 0.5 * (IM(a, i, j) + IM(a, j, i));
    }
  const double trim = (IM(t, 0, 0) + IM(t, 1, 1) + IM(t, 2, 2)) / 3.0;
  for (int i = 0; i < 3; i++) IM(t, i, i) -= trim;
  memcpy(r, t, sizeof t);
}

void or_asqtad_force_multi(const int L[4], double *const gauge[4], double *force[4],
                           const or_asqtad_coeffs_t *coef, const double eps[],
                           double *const x[], int nterms) {
  geom_t g;
  geom_init(&g, L);
  ws_t w;
  w.g = &g;
  for (int i = 0; i < 12; i++) w.t[i] = fnew(&g);
  double *mid[8] = {0}, *deriv[4];
  for (int i = 0; i < 4; i++) {
    mid[i] = fnew(&g);
    deriv[i] = fnew(&g);
  }
  // mid with the extra factor -1, odd sites negated (:491-503)
  get_mid(&g, mid, 1, eps, -1.0, x, nterms);
  if (coef->naik != 0) {
    for (int i = 4; i < 8; i++) mid[i] = fnew(&g);
    get_mid(&g, mid + 4, 3, eps, -1.0, x, nterms);
  }
  const long Vh = g.V / 2;
  for (int d = 0; d < (coef->naik != 0 ? 8 : 4); d++)
    for (long i = Vh; i < g.V; i++)
      for (int e = 0; e < 18; e++) mid[d][18 * i + e] = -mid[d][18 * i + e];
  fat_deriv(&w, gauge, deriv, coef, mid);
  if (coef->naik != 0) naik_deriv(&w, gauge, deriv, coef->naik, mid + 4);
  // force += -2 antiherm(U deriv^+) (:531-539)
  for (int mu = 0; mu < 4; mu++) {
#pragma omp parallel for
    for (long i = 0; i < g.V; i++) {
      double m[18], a[18];
      smul(m, gauge[mu] + 18 * i, 0, deriv[mu] + 18 * i, 1, 1.0, 0);
      antiherm(a, m);
      for (int e = 0; e < 18; e++) force[mu][18 * i + e] += -2.0 * a[e];
    }
  }
  for (int i = 0; i < 8; i++) free(mid[i]);
  for (int i = 0; i < 4; i++) free(deriv[i]);
  for (int i = 0; i < 12; i++) free(w.t[i]);
  geom_free(&g);
}

/* HISQ fermion force, fnmat form (lib/hisq_force_fnmat_p.c:7-231).
 * Level 2: for each epsilon-Naik group k the mids (factor +1, odd sites
 * negated, group 0 over all terms) are differentiated through the asqtad
 * smearing of W (group 0: level-2 coefficients; k > 0: the difference set)
 * and summed with weight 1 (k = 0) or eps_naik[k] into F2 (:56-118).  The
 * U(3) projection's derivative turns F2 into the mid of level 1
 * (or_u3_deriv_contract, then the adjoint: :141-152); level 1 is the fat7
 * derivative on U; deriv = 2 x level 1 and force += antiherm(U deriv^+)
 * (:171-181, 211-219). */
int or_hisq_force_multi(const int L[4], double *const U[4], double *force[4],
                        const or_asqtad_coeffs_t *fat7, const or_asqtad_coeffs_t *lev2,
                        double diff_one_link, double diff_naik, int n_naiks,
                        const double eps_naik[], const double eps[], double *const x[],
                        const int n_orders_naik[], double force_filter) {
  geom_t g;
  geom_init(&g, L);
  ws_t w;
  w.g = &g;
  for (int i = 0; i < 12; i++) w.t[i] = fnew(&g);
  const long V = g.V, Vh = V / 2;
  double *Vl[4], *Wl[4], *F1[4], *F2[4], *mid[8];
  for (int i = 0; i < 4; i++) {
    Vl[i] = fnew(&g);
    Wl[i] = fnew(&g);
    F1[i] = fnew(&g);
    F2[i] = fnew(&g);
  }
  for (int i = 0; i < 8; i++) mid[i] = fnew(&g);
  /* V = fat7(U) (user normalisation: 2 x internal link), W = U(3) projection */
  or_asqtad_coeffs_t c7 = *fat7;
  c7.naik = 0;
  or_make_imp_links(L, &c7, U, U, Vl, NULL);
  or_reunit_opts_t ro = {1, 0, 1e-8, 1e-8};
#pragma omp parallel for
  for (long i = 0; i < 4 * V; i++) {
    const int mu = (int)(i / V);
    const long s = i % V;
    double fl = 0;
    or_u3_un_analytic(Vl[mu] + 18 * s, Wl[mu] + 18 * s, &ro, &fl);
  }
  int nterms = 0;
  for (int k = 0; k < n_naiks; k++) nterms += n_orders_naik[k];
  int shift = 0;
  for (int k = 0; k < n_naiks; k++) {
    const int cnt = k == 0 ? nterms : n_orders_naik[k];
    get_mid(&g, mid, 1, eps + shift, 1.0, x + shift, cnt);
    get_mid(&g, mid + 4, 3, eps + shift, 1.0, x + shift, cnt);
    for (int d = 0; d < 8; d++)
      for (long i = Vh; i < V; i++)
        for (int e = 0; e < 18; e++) mid[d][18 * i + e] = -mid[d][18 * i + e];
    or_asqtad_coeffs_t c2 = *lev2;
    if (k > 0) {
      c2.one_link = diff_one_link;
      c2.three_staple = c2.five_staple = c2.seven_staple = c2.lepage = 0;
      c2.naik = diff_naik;
    }
    for (int mu = 0; mu < 4; mu++) fzero(&g, F1[mu]);
    fat_deriv(&w, Wl, F1, &c2, mid);
    if (c2.naik != 0) naik_deriv(&w, Wl, F1, c2.naik, mid + 4);
    const double cm = k == 0 ? 1.0 : eps_naik[k];
    for (int mu = 0; mu < 4; mu++) faxpy(&g, F2[mu], cm, F1[mu], 0);
    shift += n_orders_naik[k];
  }
  /* projection derivative; the level-1 mid is its adjoint */
  int nfilt = 0;
#pragma omp parallel for reduction(+ : nfilt)
  for (long i = 0; i < 4 * V; i++) {
    const int mu = (int)(i / V);
    const long s = i % V;
    double y[18];
    nfilt += or_u3_deriv_contract(Vl[mu] + 18 * s, F2[mu] + 18 * s, force_filter, y);
    double *m = mid[mu] + 18 * s;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        RE(m, r, c) = RE(y, c, r);
        IM(m, r, c) = -IM(y, c, r);
      }
  }
  for (int mu = 0; mu < 4; mu++) fzero(&g, F1[mu]);
  c7.naik = 0;
  fat_deriv(&w, U, F1, &c7, mid);
  /* force += antiherm(U (2 F1)^+) */
  for (int mu = 0; mu < 4; mu++) {
#pragma omp parallel for
    for (long i = 0; i < V; i++) {
      double m[18], a[18];
      smul(m, U[mu] + 18 * i, 0, F1[mu] + 18 * i, 1, 2.0, 0);
      antiherm(a, m);
      for (int e = 0; e < 18; e++) force[mu][18 * i + e] += a[e];
    }
  }
  for (int i = 0; i < 4; i++) {
    free(Vl[i]);
    free(Wl[i]);
    free(F1[i]);
    free(F2[i]);
  }
  for (int i = 0; i < 8; i++) free(mid[i]);
  for (int i = 0; i < 12; i++) free(w.t[i]);
  geom_free(&g);
  return nfilt;
}
