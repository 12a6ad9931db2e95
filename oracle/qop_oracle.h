/*
 * CPU oracle for the asqtad staggered hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C FP64 restatement of the reference algorithms, used only
 * as the checker by tests/, by the smoke check and by the cpu_baseline leg of
 * bench.py.  Nothing in the product library links or calls it.
 *
 * All fields are full-lattice arrays in the QOP "raw" layout:
 *   vector : double[V][3][2]          (colour, re/im)
 *   matrix : double[V][3][3][2]       (row, column, re/im)
 *   site index: l = ((x0*L1 + x1)*L2 + x2)*L3 + x3, even sites first:
 *   idx = l/2 (even), Vh + l/2 (odd)   (lib/layout_user.c:237-258)
 *
 * Every function cites the reference file:line it follows.  Parity pins
 * for each function live in tests/test_oracle.py (closed forms, invariants,
 * an independent second stencil restatement and brute force); see the header
 * of qop_oracle.c for the list.
 */
#ifndef QOP_ORACLE_H
#define QOP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { OR_EVEN = 0, OR_ODD = 1, OR_EVENODD = 2 };

typedef struct {
  double one_link, three_staple, five_staple, seven_staple, lepage, naik;
} or_asqtad_coeffs_t;

/* Lattice geometry (L = {L0,L1,L2,L3}, all even). */
long or_volume(const int L[4]);
long or_raw_index(const int L[4], const int x[4]);
void or_raw_coords(const int L[4], long idx, int x[4]);

/* Rephase gauge links in place (lib/common_p.c:464-526). bc = 4 complex
 * (re,im) pairs, signmask = 4 ints, r0 = 4 ints. */
void or_rephase_G(const int L[4], double *U[4], const int r0[4],
                  const double bc[8], const int signmask[4]);

/* Asqtad fattening (lib/asqtad_dslash_p.c:458-633).  lng may be NULL. */
void or_make_imp_links(const int L[4], const or_asqtad_coeffs_t *c,
                       double *const gf[4], double *const gfLong[4],
                       double *fat[4], double *lng[4]);

/* Plaquette, normalised as examples/test_common.c:171-228. */
double or_plaquette(const int L[4], double *const U[4]);

/* Link object: user-normalised fat/long (lng may be NULL) -> internal
 * links scaled by 0.5 and double stored (asqtad_dslash_p.c:79-133,152-223). */
typedef struct or_links_s or_links_t;
or_links_t *or_links_create(const int L[4], double *const fat[4],
                            double *const lng[4]);
void or_links_destroy(or_links_t *l);

/* QOP_asqtad_dslash_qdp semantics (asqtad_dslash_p.c:1048-1092). */
void or_dslash(or_links_t *l, double mass, double *out, const double *in,
               int eo_out, int eo_in);
/* QOP_asqtad_dslash_dir (asqtad_dslash_p.c:1140-1377): out += one-direction
 * transport on eo_out; bsign multiplies the backward terms (+1 style 1,
 * -1 style 0). */
void or_dslash_dir(or_links_t *l, int dir, int fb, double wtfat, double wtlong, double *out,
                   const double *in, int eo_out, double bsign);
/* FP32 storage of links and of every vector the operator and the solvers
 * store (the single-precision build, lib/pcstubs/asqtad_invert_f3.c). */
void or_links_set_single(or_links_t *l, int on);
/* Independent restatement of the full operator from the user links,
 * following examples/asqtad_test.c:37-81 (check_op). Writes all sites. */
void or_dslash_check_op(const int L[4], double *const fat[4],
                        double *const lng[4], double mass, double *out,
                        const double *in);
/* QOP_asqtad_ddagd_norm2 (asqtad_dslash_p.c:1121-1138). */
double or_ddagd_norm2(or_links_t *l, double mass, double *out,
                      const double *in, int eo);

typedef struct {
  double mixed_rsq;
  int max_iter, restart, max_restarts, evenodd;
} or_invert_arg_t;
typedef struct {
  double rsqmin, final_rsq, relmin, final_rel;
  int final_iter, final_restart;
} or_resid_arg_t;

/* QOP_asqtad_invert (asqtad_invert_p.c:134-144, 649-851). */
void or_invert(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t *ra,
               double mass, double *out, const double *in,
               double *final_flop);
/* QOP_asqtad_invert_multi (asqtad_invert_p.c:853-1059).  masses, out, ra
 * are flattened [src][mass] with nmass[src] entries per source. */
void or_invert_multi(or_links_t *l, or_invert_arg_t *ia, or_resid_arg_t *ra,
                     const double *masses, const int *nmass, double **out,
                     double *const *in, int nsrc, double *final_flop);
/* eigCG state (the reference keeps it in the links object:
 * asqtad_invert_p.c:683-687, invert_eigcg_p.c:618-1068). */
typedef struct or_eigcg_s or_eigcg_t;
or_eigcg_t *or_eigcg_create(void);
void or_eigcg_destroy(or_eigcg_t *eg);
void or_eigcg_set(or_eigcg_t *eg, int numax, int m, int nev);
/* number of deflation vectors nu; their Ritz values into l[0:min(nu, maxn)] */
int or_eigcg_nu(const or_eigcg_t *eg, double *l, int maxn);
/* QOP_asqtad_invert (EVENODD) with cgtype = 1: eigCG inner solves. */
void or_invert_eigcg(or_links_t *l, or_eigcg_t *eg, or_invert_arg_t *ia, or_resid_arg_t *ra,
                     double mass, double *out, const double *in, double *final_flop);
/* Fermilab relative residual (invert_utilities_p.c:7-40). */
double or_relnorm2(or_links_t *l, const double *r, const double *x, int eo);

/* HISQ link preparation (hisq_oracle.c). */
typedef struct {
  int allow_svd, svd_only;
  double svd_rel_error, svd_abs_error;
} or_reunit_opts_t;
/* W = V (V^dagger V)^-1/2 by the analytic cubic
 * (hisq_reunit_utilities_p.c:97-297); returns 1 when the fallback (unitary
 * polar factor from a Jacobi SVD) was taken; flops accumulated. */
int or_u3_un_analytic(const double *V, double *W, const or_reunit_opts_t *o, double *flops);
void or_u3_polar_jacobi(const double *V, double *W);
/* fat7 smearing of U then U(3) projection of every link into W
 * (hisq_links_p.c:62-152); returns the number of fallback projections. */
int or_hisq_W(const int L[4], const or_asqtad_coeffs_t *fat7, double *const U[4], double *W[4],
              const or_reunit_opts_t *o, double *flops);

/* Derivative of the U(3) projection W = V (V^+V [+ f])^-1/2 contracted with
 * F: Y(n,m) = d[tr(W F^+) + tr(W^+ F)]/dV_mn (lib/hisq_reunit1_p.c:366-433);
 * returns 1 when the force filter f acted. */
int or_u3_deriv_contract(const double *V, const double *F, double filter, double *Y);
/* HISQ fermion force, fnmat form (lib/hisq_force_fnmat_p.c:7-231):
 * force[mu] += antiherm(U_mu deriv_mu^+).  fat7: level-1 coefficients (naik
 * ignored, lepage as the caller wants it), lev2: level-2 asqtad
 * coefficients, diff_*: the epsilon-Naik difference set; the x are grouped
 * by n_orders_naik[0..n_naiks-1].  Returns the number of filtered links. */
int or_hisq_force_multi(const int L[4], double *const U[4], double *force[4],
                        const or_asqtad_coeffs_t *fat7, const or_asqtad_coeffs_t *lev2,
                        double diff_one_link, double diff_naik, int n_naiks,
                        const double eps_naik[], const double eps[], double *const x[],
                        const int n_orders_naik[], double force_filter);

/* Number of iterations done by the last inner solver calls (diagnostics). */
/* Asqtad fermion force, fnmat form (lib/asqtad_force_fnmat_p.c:474-546):
 * force[mu] += -2 antiherm(U_mu deriv_mu^+), deriv the derivative of the fat
 * and Naik links of sum_i eps_i x_i x_i^+ terms (force_oracle.c). */
void or_asqtad_force_multi(const int L[4], double *const gauge[4], double *force[4],
                           const or_asqtad_coeffs_t *coef, const double eps[],
                           double *const x[], int nterms);

int or_set_verbose(int v);
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
