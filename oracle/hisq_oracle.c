/*
 * CPU oracle for the HISQ link preparation -- TEST INFRASTRUCTURE ONLY.
 *
 * Function -> reference file:line it follows -> what pins it (tests/):
 *   or_u3_un_analytic   lib/hisq_reunit_utilities_p.c:97-297   W^dagger W = 1 (1e-14 on fat7 links,
 *                                                              1e-11 on generic V) and W^dagger V
 *                                                              hermitian positive (the polar factor);
 *                                                              agreement with the Jacobi path below and
 *                                                              with numpy's SVD on random V; V = c U
 *                                                              (U unitary) gives W = U
 *   or_u3_polar_jacobi  (fallback of the same function, :197-245: the reference
 *                       takes W = U_left V_right^dagger from its own SVD routine;
 *                       this computes the same matrix from a one-sided Jacobi SVD)
 *                                                              numpy SVD on random and
 *                                                              ill-conditioned V
 *   or_hisq_W           lib/hisq_links_p.c:62-152 (fat7 stage, x2 of the
 *                       internal 1/2, U(3) projection)            free field: fat7 weight 1 so W = U;
 *                                                              gauge covariance of W
 *   or_u3_deriv_contract lib/hisq_reunit1_p.c:366-433 (+ the derivative of
 *                       hisq_reunit_utilities_p.c:301-640, restated from the
 *                       mathematics in the eigenbasis of V^dagger V)
 *                                                              central differences of
 *                                                              tr(W F^+) + tr(W^+ F) on
 *                                                              generic and near-unitary V,
 *                                                              with and without the filter
 * The level-2 asqtad smearing of W is or_make_imp_links (qop_oracle.c), as
 * lib/hisq_links_p.c:233 calls QOP_asqtad_create_L_from_G on W.
 */
#include "qop_oracle.h"

#include <math.h>
#include <string.h>

#define RE(m, i, j) (m)[6 * (i) + 2 * (j)]
#define IM(m, i, j) (m)[6 * (i) + 2 * (j) + 1]

/* c = a*b */
static void mm(double *c, const double *a, const double *b) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, i, k) * RE(b, k, j) - IM(a, i, k) * IM(b, k, j);
        im += RE(a, i, k) * IM(b, k, j) + IM(a, i, k) * RE(b, k, j);
      }
      RE(t, i, j) = re;
      IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}

/* c = a^dagger b */
static void mam(double *c, const double *a, const double *b) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, k, i) * RE(b, k, j) + IM(a, k, i) * IM(b, k, j);
        im += RE(a, k, i) * IM(b, k, j) - IM(a, k, i) * RE(b, k, j);
      }
      RE(t, i, j) = re;
      IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}

/* |det V|^2 */
static double det_abs2(const double *v) {
  /* det = sum over the first row of cofactors */
  double dr = 0, di = 0;
  for (int j = 0; j < 3; j++) {
    int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    /* minor m = v[1][j1] v[2][j2] - v[1][j2] v[2][j1] */
    double mr = RE(v, 1, j1) * RE(v, 2, j2) - IM(v, 1, j1) * IM(v, 2, j2) -
                (RE(v, 1, j2) * RE(v, 2, j1) - IM(v, 1, j2) * IM(v, 2, j1));
    double mi = RE(v, 1, j1) * IM(v, 2, j2) + IM(v, 1, j1) * RE(v, 2, j2) -
                (RE(v, 1, j2) * IM(v, 2, j1) + IM(v, 1, j2) * RE(v, 2, j1));
    dr += RE(v, 0, j) * mr - IM(v, 0, j) * mi;
    di += RE(v, 0, j) * mi + IM(v, 0, j) * mr;
  }
  return dr * dr + di * di;
}

/* Unitary polar factor W = U_left V_right^dagger of V from a one-sided
 * (Hestenes) Jacobi SVD: complex plane rotations G applied to the columns of
 * A = V until they are mutually orthogonal, A = V R with R = prod G unitary;
 * then A = U diag(sigma), V = U diag(sigma) R^dagger and W = U R^dagger.  A
 * zero column of A (singular V) takes the unit vector completing U. */
/* One-sided Jacobi: a = V r with orthogonal columns, r unitary (the right
 * singular vectors of V, i.e. the eigenvectors of V^dagger V). */
static void jacobi_right(const double *V, double *a, double *r) {
  memcpy(a, V, 18 * sizeof(double));
  memset(r, 0, 18 * sizeof(double));
  for (int i = 0; i < 3; i++) RE(r, i, i) = 1;
  for (int sweep = 0; sweep < 60; sweep++) {
    double worst = 0;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double al = 0, be = 0, gr = 0, gi = 0;
        for (int k = 0; k < 3; k++) {
          al += RE(a, k, p) * RE(a, k, p) + IM(a, k, p) * IM(a, k, p);
          be += RE(a, k, q) * RE(a, k, q) + IM(a, k, q) * IM(a, k, q);
          /* gamma = a_p^dagger a_q */
          gr += RE(a, k, p) * RE(a, k, q) + IM(a, k, p) * IM(a, k, q);
          gi += RE(a, k, p) * IM(a, k, q) - IM(a, k, p) * RE(a, k, q);
        }
        double g = sqrt(gr * gr + gi * gi);
        if (g == 0 || g <= 1e-17 * sqrt(al * be)) continue;
        double rel = g / sqrt(al * be);
        if (rel > worst) worst = rel;
        double cph = gr / g, sph = gi / g;
        double zeta = (be - al) / (2 * g);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        double c = 1 / sqrt(1 + t * t), s = c * t;
        /* columns p, q of a and r:  p' = c p - s e^{-i phi} q,  q' = s e^{i phi} p + c q */
        for (int m = 0; m < 2; m++) {
          double *x = m ? r : a;
          for (int k = 0; k < 3; k++) {
            double pr = RE(x, k, p), pi = IM(x, k, p), qr = RE(x, k, q), qi = IM(x, k, q);
            RE(x, k, p) = c * pr - s * (cph * qr + sph * qi);
            IM(x, k, p) = c * pi - s * (cph * qi - sph * qr);
            RE(x, k, q) = s * (cph * pr - sph * pi) + c * qr;
            IM(x, k, q) = s * (cph * pi + sph * pr) + c * qi;
          }
        }
      }
    if (worst < 1e-16) break;
  }
}

void or_u3_polar_jacobi(const double *V, double *W) {
  double a[18], r[18];
  jacobi_right(V, a, r);
  /* U: normalised columns of a */
  double u[18], sig[3], smax = 0;
  for (int j = 0; j < 3; j++) {
    double n = 0;
    for (int k = 0; k < 3; k++) n += RE(a, k, j) * RE(a, k, j) + IM(a, k, j) * IM(a, k, j);
    sig[j] = sqrt(n);
    if (sig[j] > smax) smax = sig[j];
  }
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) {
      double f = sig[j] > 1e-300 ? 1 / sig[j] : 0;
      RE(u, k, j) = RE(a, k, j) * f;
      IM(u, k, j) = IM(a, k, j) * f;
    }
  for (int j = 0; j < 3; j++) {
    if (sig[j] > 1e-14 * smax) continue;
    /* complete: u_j = conj(u_j1 x u_j2) */
    int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    for (int k = 0; k < 3; k++) {
      int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      double xr = RE(u, k1, j1) * RE(u, k2, j2) - IM(u, k1, j1) * IM(u, k2, j2) -
                  (RE(u, k2, j1) * RE(u, k1, j2) - IM(u, k2, j1) * IM(u, k1, j2));
      double xi = RE(u, k1, j1) * IM(u, k2, j2) + IM(u, k1, j1) * RE(u, k2, j2) -
                  (RE(u, k2, j1) * IM(u, k1, j2) + IM(u, k2, j1) * RE(u, k1, j2));
      RE(u, k, j) = xr;
      IM(u, k, j) = -xi;
    }
  }
  /* W = U R^dagger */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(u, i, k) * RE(r, j, k) + IM(u, i, k) * IM(r, j, k);
        im += IM(u, i, k) * RE(r, j, k) - RE(u, i, k) * IM(r, j, k);
      }
      RE(W, i, j) = re;
      IM(W, i, j) = im;
    }
}

int or_u3_un_analytic(const double *V, double *W, const or_reunit_opts_t *o, double *flops) {
  const double eps = 1.0e-14; /* QOP_U3_UNIT_ANALYTIC_EPS in double precision */
  double nfl = 0, det_check = 0;
  int svd = 0;
  if (o->allow_svd) det_check = det_abs2(V);
  nfl += 67;
  double q[18], q2[18], q3[18];
  mam(q, V, V);
  mm(q2, q, q);
  mm(q3, q2, q);
  nfl += 3 * 198;
  double c0 = RE(q, 0, 0) + RE(q, 1, 1) + RE(q, 2, 2);
  double c1 = (RE(q2, 0, 0) + RE(q2, 1, 1) + RE(q2, 2, 2)) / 2;
  double c2 = (RE(q3, 0, 0) + RE(q3, 1, 1) + RE(q3, 2, 2)) / 3;
  double S = c1 / 3 - c0 * (c0 / 18);
  double g0, g1, g2;
  nfl += 11;
  if (fabs(S) < eps) {
    g0 = g1 = g2 = c0 / 3;
    nfl += 3;
  } else {
    double R = c2 / 2 - c0 * (c1 / 3) + c0 * c0 * (c0 / 27);
    S = sqrt(S);
    double S3 = S * S * S, RoS = R / S3, theta;
    nfl += 15;
    if (!(fabs(RoS) < 1.0)) theta = R > 0 ? 0.0 : M_PI;
    else theta = acos(RoS);
    double th3 = theta / 3, p23 = 2 * M_PI / 3;
    g0 = c0 / 3 + 2 * S * cos(th3);
    g1 = c0 / 3 + 2 * S * cos(th3 + p23);
    g2 = c0 / 3 + 2 * S * cos(th3 + 2 * p23);
    nfl += 21;
  }
  if (o->allow_svd) {
    if (!o->svd_only) {
      if (det_check != 0 && fabs(det_check - g0 * g1 * g2) / fabs(det_check) > o->svd_rel_error) {
        svd = 1;
        nfl += 4;
      }
      if (det_check < o->svd_abs_error) svd = 1;
    } else {
      svd = 1;
    }
  }
  if (svd) {
    or_u3_polar_jacobi(V, W);
    nfl += 198;
  } else {
    double g0s = sqrt(g0), g1s = sqrt(g1), g2s = sqrt(g2);
    double us = g1s + g2s, ws = g1s * g2s, vs = g0s * us + ws;
    us += g0s;
    ws *= g0s;
    double den = ws * (us * vs - ws);
    double f0 = (us * vs * vs - ws * (us * us + vs)) / den;
    double f1 = (2 * us * vs - ws - us * us * us) / den;
    double f2 = us / den;
    nfl += 30;
    double s[18];
    for (int i = 0; i < 18; i++) s[i] = f1 * q[i] + f2 * q2[i];
    for (int i = 0; i < 3; i++) RE(s, i, i) += f0;
    nfl += 57;
    mm(W, V, s);
  }
  if (flops) *flops += nfl;
  return svd;
}

int or_hisq_W(const int L[4], const or_asqtad_coeffs_t *fat7, double *const U[4], double *W[4],
              const or_reunit_opts_t *o, double *flops) {
  const long V = or_volume(L);
  double *Vl[4];
  for (int mu = 0; mu < 4; mu++) Vl[mu] = W[mu]; /* the fat7 links are projected in place */
  /* fat7 stage: asqtad smearing with Naik 0; the user-normalised fat link is
   * the reference's 2 x (internal link of 1/2) (hisq_links_p.c:118-128) */
  or_asqtad_coeffs_t c = *fat7;
  c.naik = 0;
  or_make_imp_links(L, &c, U, U, Vl, NULL);
  int nsvd = 0;
  double fl = 0;
#pragma omp parallel for reduction(+ : nsvd, fl)
  for (long i = 0; i < 4 * V; i++) {
    int mu = (int)(i / V);
    long s = i % V;
    double w[18], f = 0;
    nsvd += or_u3_un_analytic(Vl[mu] + 18 * s, w, o, &f);
    memcpy(W[mu] + 18 * s, w, sizeof w);
    fl += f;
  }
  if (flops) *flops += fl;
  return nsvd;
}

/* Derivative of the U(3) projection contracted with the force of the level
 * above (lib/hisq_reunit1_p.c:366-433 with hisq_reunit_utilities_p.c:301-640;
 * restated from the mathematics, not from that routine's closed form).
 *
 * With W = V Q^-1/2, Q = V^dagger V (+ f 1 when the force filter acts),
 * F the accumulated derivative with respect to W and
 * Phi(V) = tr(W F^dagger) + tr(W^dagger F), the reference returns
 * Y(n,m) = dPhi/dV_mn (V^* held fixed).  In the eigenbasis Q = E diag(l) E^+,
 * s = sqrt(l):  dQ^-1/2 = E [G o (E^+ dQ E)] E^+, G_ij = -1 / (s_i s_j (s_i + s_j)),
 * which gives
 *   Y = Q^-1/2 F^+ + (N + N^+) V^+,   N = E [G o (E^+ F^+ V E)] E^+.
 * The filter (hisq_reunit_utilities_p.c:437-456): when the smallest
 * eigenvalue is below f > 0, f is added to all of them.  Returns 1 when the
 * filter acted. */
static void mmul_t(double *c, const double *a, int ta, const double *b, int tb) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        double ar = ta ? RE(a, k, i) : RE(a, i, k), ai = ta ? -IM(a, k, i) : IM(a, i, k);
        double br = tb ? RE(b, j, k) : RE(b, k, j), bi = tb ? -IM(b, j, k) : IM(b, k, j);
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      RE(t, i, j) = re;
      IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}

int or_u3_deriv_contract(const double *V, const double *F, double filter, double *Y) {
  double a[18], E[18], l[3], s[3];
  jacobi_right(V, a, E);
  for (int j = 0; j < 3; j++) {
    l[j] = 0;
    for (int k = 0; k < 3; k++) l[j] += RE(a, k, j) * RE(a, k, j) + IM(a, k, j) * IM(a, k, j);
  }
  int filtered = 0;
  if (filter > 0 && (l[0] < filter || l[1] < filter || l[2] < filter)) {
    for (int j = 0; j < 3; j++) l[j] += filter;
    filtered = 1;
  }
  for (int j = 0; j < 3; j++) s[j] = sqrt(l[j]);
  /* Q^-1/2 = E diag(1/s) E^+ */
  double D[18], Qm[18], t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      RE(t, i, j) = RE(E, i, j) / s[j];
      IM(t, i, j) = IM(E, i, j) / s[j];
    }
  mmul_t(Qm, t, 0, E, 1);
  /* M' = E^+ F^+ V E, N' = G o M', N = E N' E^+ */
  double M[18], Mp[18], N[18];
  mmul_t(M, F, 1, V, 0);
  mmul_t(t, E, 1, M, 0);
  mmul_t(Mp, t, 0, E, 0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double g = -1.0 / (s[i] * s[j] * (s[i] + s[j]));
      RE(Mp, i, j) *= g;
      IM(Mp, i, j) *= g;
    }
  mmul_t(t, E, 0, Mp, 0);
  mmul_t(N, t, 0, E, 1);
  /* Y = Q^-1/2 F^+ + (N + N^+) V^+ */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      RE(D, i, j) = RE(N, i, j) + RE(N, j, i);
      IM(D, i, j) = IM(N, i, j) - IM(N, j, i);
    }
  double y1[18], y2[18];
  mmul_t(y1, Qm, 0, F, 1);
  mmul_t(y2, D, 0, V, 1);
  for (int e = 0; e < 18; e++) Y[e] = y1[e] + y2[e];
  return filtered;
}
