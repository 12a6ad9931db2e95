/*
 * qop_mi355x.h -- C ABI of the MI355X-native asqtad hot path.
 *
 * Drop-in for the asqtad subset of the QOPQDP C API (Nc = 3, double "D3"
 * and single "F3" precision).  The type definitions below are laid out
 * exactly as in the reference's include/qop_int.h so code compiled against
 * QOPQDP's headers passes the same structs; the functions carry the same
 * names, arguments and conventions as the reference prototypes in
 * include/qop_pc.h (instantiated by include/qop_d3.h / qop_f3.h).
 *
 * Conventions kept from the reference:
 *  - fields in the "raw" layout: vector QOP_Real[V][3][2], link matrix
 *    QOP_Real[V][3][3][2] (row major), site index = lexicographic
 *    ((x0*L1+x1)*L2+x2)*L3+x3 halved, even sites first
 *    (lib/layout_user.c:237-258);
 *  - fermion links are stored internally multiplied by 1/2
 *    (lib/asqtad_dslash_p.c:201-206);
 *  - an EVEN (ODD) subset solve returns out = m A^-1 in with
 *    A = m^2 - D_eo D_oe, and res_arg->final_rsq is |r|^2/|in|^2
 *    (README:24-48, lib/asqtad_invert_p.c:1-40, 603-610).
 *
 * Device memory exhaustion is not fatal: an entry point with a QOP_info_t
 * argument sets info->status = QOP_MEM_ERROR (include/qop_int.h:19) and
 * returns without a result, the create_* functions return NULL, and the
 * memory they had taken is given back; other failures print the HIP error
 * and abort, as the reference exits on its fatal conditions.
 *
 * The QDP-typed entry points (*_qdp) of the reference are not provided:
 * there is no QDP/C here.  All compute runs on the GPU (hand-written HIP
 * kernels for gfx950); there is no CPU path.
 *
 * Entry points named QOPX_* are extensions of this library with no
 * counterpart in the reference: seeded inputs, batched mixed-precision
 * solves, a continuous-batching request loop and timing helpers.
 */
#ifndef QOP_MI355X_H
#define QOP_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- types: same layout as include/qop_int.h:1-186 ------------------- */
#define QOP_VERB_OFF 0
#define QOP_VERB_LOW 1
#define QOP_VERB_MED 2
#define QOP_VERB_HI 3
#define QOP_VERB_DEBUG 4

typedef enum {
  QOP_SUCCESS = 0,
  QOP_FAIL = 1,
  QOP_BC_ERROR,
  QOP_MEM_ERROR,
  QOP_TOP_ERROR
} QOP_status_t; /* include/qop_int.h:15-21 */

typedef enum { QOP_EVEN = 0, QOP_ODD = 1, QOP_EVENODD = 2 } QOP_evenodd_t; /* :23-27 */

typedef float QOP_F_Real;
typedef double QOP_D_Real;

typedef struct {
  int (*node_number)(const int coords[]);
  int (*node_index)(const int coords[]);
  int latdim;
  int *latsize;
  int machdim;
  int *machsize;
  int this_node;
  int sites_on_node;
} QOP_layout_t; /* :51-60 */
#define QOP_LAYOUT_ZERO ((QOP_layout_t){NULL, NULL, 0, NULL, 0, NULL, 0, 0})

typedef struct {
  double re;
  double im;
} QOP_Complex;

typedef struct {
  QOP_Complex *phase;
} QOP_bc_t; /* :70-72 */

typedef struct {
  int *signmask;
} QOP_staggered_sign_t; /* :78-80 */

typedef struct {
  char *tag;
  double value;
  void *extra;
} QOP_opt_t; /* :79-83 */

typedef struct {
  double final_sec;
  double final_flop;
  QOP_status_t status;
  int count1, count2;
} QOP_info_t; /* :88-94 */
#define QOP_INFO_ZERO ((QOP_info_t){0, 0, QOP_SUCCESS, 0, 0})

typedef struct {
  double mixed_rsq;
  int max_iter;
  int restart;
  int max_restarts;
  QOP_evenodd_t evenodd;
} QOP_invert_arg_t; /* :118-125 */
#define QOP_INVERT_ARG_DEFAULT ((QOP_invert_arg_t){1e9, 2000, 1000, 5, QOP_EVENODD})

typedef struct {
  double rsqmin;
  double final_rsq;
  double relmin;
  double final_rel;
  int final_iter;
  int final_restart;
} QOP_resid_arg_t; /* :128-136 */
#define QOP_RESID_ARG_DEFAULT ((QOP_resid_arg_t){1e-6, 0, 0, 0, 0, 0})

typedef struct {
  double one_link;
  double three_staple;
  double five_staple;
  double seven_staple;
  double lepage;
  double naik;
} QOP_asqtad_coeffs_t; /* :150-157 */
#define QOP_ASQTAD_COEFFS_ZERO ((QOP_asqtad_coeffs_t){0, 0, 0, 0, 0, 0})

typedef enum { QOP_UNITARIZE_U3 = 0, QOP_UNITARIZE_SU3 = 1 } QOP_hisq_unitarize_group_t; /* :29-32 */
typedef enum {
  QOP_UNITARIZE_NONE = 0,
  QOP_UNITARIZE_RATIONAL = 3,
  QOP_UNITARIZE_ANALYTIC = 5
} QOP_hisq_unitarize_method_t; /* :35-43 */
#define QOP_MAX_NAIK 20
#define QOP_EPS_NAIK_ZERO {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}
typedef struct {
  int n_naiks;
  double eps_naik[QOP_MAX_NAIK];
  QOP_hisq_unitarize_group_t ugroup;
  QOP_hisq_unitarize_method_t umethod;
  double fat7_one_link;
  double fat7_three_staple;
  double fat7_five_staple;
  double fat7_seven_staple;
  double fat7_lepage;
  double asqtad_one_link;
  double asqtad_three_staple;
  double asqtad_five_staple;
  double asqtad_seven_staple;
  double asqtad_lepage;
  double asqtad_naik;
  double difference_one_link;
  double difference_naik;
} QOP_hisq_coeffs_t; /* :166-184 */
#define QOP_HISQ_COEFFS_ZERO                                                   \
  ((QOP_hisq_coeffs_t){1, QOP_EPS_NAIK_ZERO, QOP_UNITARIZE_U3, QOP_UNITARIZE_RATIONAL, \
                       0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0})
/* U(3) projections that fell back from the analytic formula (:115) */
#define QOP_info_hisq_svd_counter(info) ((info)->count1)

/* opaque handles (include/qop_pc.h:10-18) */
typedef struct QOP_D3_ColorVector_struct QOP_D3_ColorVector;
typedef struct QOP_F3_ColorVector_struct QOP_F3_ColorVector;
typedef struct QOP_D3_GaugeField_struct QOP_D3_GaugeField;
typedef struct QOP_F3_GaugeField_struct QOP_F3_GaugeField;
typedef struct QOP_D3_FermionLinksAsqtad_struct QOP_D3_FermionLinksAsqtad;
typedef struct QOP_F3_FermionLinksAsqtad_struct QOP_F3_FermionLinksAsqtad;
typedef struct QOP_D3_FermionLinksHisq_struct QOP_D3_FermionLinksHisq;
typedef struct QOP_F3_FermionLinksHisq_struct QOP_F3_FermionLinksHisq;
typedef struct QOP_D3_Force_struct QOP_D3_Force;
typedef struct QOP_F3_Force_struct QOP_F3_Force;

/* ---- general routines (include/qop_int.h:230-251; lib/common.c) ------ */
QOP_status_t QOP_init(QOP_layout_t *layout);         /* lib/common.c:40-83 */
QOP_status_t QOP_finalize(void);                     /* lib/common.c:86-94 */
int QOP_is_initialized(void);                        /* lib/common.c:97-100 */
const char *QOP_version_str(void);                   /* lib/common.c:105-108 */
int QOP_version_int(void);                           /* lib/common.c:111-116 */
int QOP_verbose(int level);                          /* lib/common.c:119-124 */
int QOP_profcontrol(int level);                      /* lib/common.c:127-132 */
int QOP_node_number_raw(int coords[]);               /* lib/common.c:135-139 */
int QOP_node_index_raw_V(int coords[], QOP_evenodd_t evenodd); /* lib/common.c:142-146 */
int QOP_node_index_raw_G(int coords[], QOP_evenodd_t evenodd); /* lib/common.c:156-160 */
int QOP_sites_on_node_raw_V(QOP_evenodd_t evenodd);  /* lib/common.c:170-173 */
int QOP_sites_on_node_raw_G(QOP_evenodd_t evenodd);  /* lib/common.c:182-185 */
/* option tags st, ns, nm, cg, eigcg_* as lib/asqtad_invert.c:68-112; tags
 * starting "qmx_" select this library's solvers (see QOPX options). */
QOP_status_t QOP_asqtad_invert_set_opts(QOP_opt_t opts[], int nopts);

/* ---- fields (include/qop_pc.h:49-121; lib/common_p.c) ---------------- */
QOP_D3_ColorVector *QOP_D3_create_V_from_raw(QOP_D_Real *src, QOP_evenodd_t evenodd); /* common_p.c:37-48 */
QOP_F3_ColorVector *QOP_F3_create_V_from_raw(QOP_F_Real *src, QOP_evenodd_t evenodd);
void QOP_D3_extract_V_to_raw(QOP_D_Real *dest, QOP_D3_ColorVector *src, QOP_evenodd_t evenodd); /* common_p.c:100-104 */
void QOP_F3_extract_V_to_raw(QOP_F_Real *dest, QOP_F3_ColorVector *src, QOP_evenodd_t evenodd);
void QOP_D3_destroy_V(QOP_D3_ColorVector *field); /* common_p.c:133-145 */
void QOP_F3_destroy_V(QOP_F3_ColorVector *field);
QOP_D3_ColorVector *QOP_D3_convert_V_from_raw(QOP_D_Real *src, QOP_evenodd_t evenodd); /* common_p.c:173-183 */
QOP_F3_ColorVector *QOP_F3_convert_V_from_raw(QOP_F_Real *src, QOP_evenodd_t evenodd);
QOP_D_Real *QOP_D3_convert_V_to_raw(QOP_D3_ColorVector *src, QOP_evenodd_t evenodd); /* common_p.c:224-240 */
QOP_F_Real *QOP_F3_convert_V_to_raw(QOP_F3_ColorVector *src, QOP_evenodd_t evenodd);

QOP_D3_GaugeField *QOP_D3_create_G_from_raw(QOP_D_Real *links[], QOP_evenodd_t evenodd); /* common_p.c:67-78 */
QOP_F3_GaugeField *QOP_F3_create_G_from_raw(QOP_F_Real *links[], QOP_evenodd_t evenodd);
void QOP_D3_extract_G_to_raw(QOP_D_Real *dest[], QOP_D3_GaugeField *src, QOP_evenodd_t evenodd); /* common_p.c:114-121 */
void QOP_F3_extract_G_to_raw(QOP_F_Real *dest[], QOP_F3_GaugeField *src, QOP_evenodd_t evenodd);
void QOP_D3_destroy_G(QOP_D3_GaugeField *field); /* common_p.c:147-160 */
void QOP_F3_destroy_G(QOP_F3_GaugeField *field);
/* the field keeps the caller's raw arrays until convert_G_to_raw, which
 * writes the field back into them and returns the same pointer array (an
 * array of 4 malloc'd fields when the field was not converted from raw);
 * the object is destroyed (common_p.c:199-208, 244-257) */
QOP_D3_GaugeField *QOP_D3_convert_G_from_raw(QOP_D_Real *links[], QOP_evenodd_t evenodd); /* common_p.c:199-208 */
QOP_F3_GaugeField *QOP_F3_convert_G_from_raw(QOP_F_Real *links[], QOP_evenodd_t evenodd);
QOP_D_Real **QOP_D3_convert_G_to_raw(QOP_D3_GaugeField *src, QOP_evenodd_t evenodd); /* common_p.c:244-257 */
QOP_F_Real **QOP_F3_convert_G_to_raw(QOP_F3_GaugeField *src, QOP_evenodd_t evenodd);
/* staggered signs and boundary phases in place (common_p.c:464-526) */
void QOP_D3_rephase_G(QOP_D3_GaugeField *links, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign);
void QOP_F3_rephase_G(QOP_F3_GaugeField *links, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign);

/* Force fields: one colour matrix per site and direction, raw layout as the
 * gauge field.  Sites outside evenodd are zero after create. */
QOP_D3_Force *QOP_D3_create_F_from_raw(QOP_D_Real *force[], QOP_evenodd_t evenodd); /* common_p.c:80-96 */
QOP_F3_Force *QOP_F3_create_F_from_raw(QOP_F_Real *force[], QOP_evenodd_t evenodd);
void QOP_D3_extract_F_to_raw(QOP_D_Real *dest[], QOP_D3_Force *src, QOP_evenodd_t evenodd); /* common_p.c:123-130 */
void QOP_F3_extract_F_to_raw(QOP_F_Real *dest[], QOP_F3_Force *src, QOP_evenodd_t evenodd);
void QOP_D3_destroy_F(QOP_D3_Force *field); /* common_p.c:160-169 */
void QOP_F3_destroy_F(QOP_F3_Force *field);
QOP_D3_Force *QOP_D3_convert_F_from_raw(QOP_D_Real *force[], QOP_evenodd_t evenodd); /* common_p.c:212-221 */
QOP_F3_Force *QOP_F3_convert_F_from_raw(QOP_F_Real *force[], QOP_evenodd_t evenodd);
QOP_D_Real **QOP_D3_convert_F_to_raw(QOP_D3_Force *src, QOP_evenodd_t evenodd); /* common_p.c:260-273 */
QOP_F_Real **QOP_F3_convert_F_to_raw(QOP_F3_Force *src, QOP_evenodd_t evenodd);

/* ---- asqtad links (include/qop_pc.h:217-271; lib/asqtad_dslash_p.c) -- */
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_create_L_from_raw(QOP_D_Real *fatlinks[], QOP_D_Real *longlinks[], QOP_evenodd_t evenodd); /* :307-329 */
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_create_L_from_raw(QOP_F_Real *fatlinks[], QOP_F_Real *longlinks[], QOP_evenodd_t evenodd);
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_create_L_from_G(QOP_info_t *info, QOP_asqtad_coeffs_t *coeffs, QOP_D3_GaugeField *gauge); /* :642-664 */
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_create_L_from_G(QOP_info_t *info, QOP_asqtad_coeffs_t *coeffs, QOP_F3_GaugeField *gauge);
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_create_L_from_G2(QOP_info_t *info, QOP_asqtad_coeffs_t *coeffs, QOP_D3_GaugeField *gFat, QOP_D3_GaugeField *gLong); /* :668-689 */
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_create_L_from_G2(QOP_info_t *info, QOP_asqtad_coeffs_t *coeffs, QOP_F3_GaugeField *gFat, QOP_F3_GaugeField *gLong);
void QOP_D3_asqtad_extract_L_to_raw(QOP_D_Real *fatlinks[], QOP_D_Real *longlinks[], QOP_D3_FermionLinksAsqtad *src, QOP_evenodd_t evenodd); /* :916-938 */
void QOP_F3_asqtad_extract_L_to_raw(QOP_F_Real *fatlinks[], QOP_F_Real *longlinks[], QOP_F3_FermionLinksAsqtad *src, QOP_evenodd_t evenodd);
/* The link object takes the caller's raw arrays: they hold the links x 1/2
 * from here on (the reference scales them in place, :152-223, 281-304) and
 * must not be changed or freed until convert_L_to_raw, which writes the
 * current links back into them (user normalisation, whole lattice), stores
 * the caller's pointer arrays in *fatlinks / *longlinks (malloc'd arrays of
 * malloc'd fields when the object was not converted from raw; *longlinks
 * NULL without Naik links) and destroys the object.  The reference declares
 * convert_L_to_raw (include/qop_pc.h:247-250) but defines no body. */
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_convert_L_from_raw(QOP_D_Real *fatlinks[], QOP_D_Real *longlinks[], QOP_evenodd_t evenodd); /* :281-304 */
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_convert_L_from_raw(QOP_F_Real *fatlinks[], QOP_F_Real *longlinks[], QOP_evenodd_t evenodd);
void QOP_D3_asqtad_convert_L_to_raw(QOP_D_Real ***fatlinks, QOP_D_Real ***longlinks, QOP_D3_FermionLinksAsqtad *src, QOP_evenodd_t evenodd); /* include/qop_pc.h:247-250 */
void QOP_F3_asqtad_convert_L_to_raw(QOP_F_Real ***fatlinks, QOP_F_Real ***longlinks, QOP_F3_FermionLinksAsqtad *src, QOP_evenodd_t evenodd);
void QOP_D3_asqtad_load_L_from_raw(QOP_D3_FermionLinksAsqtad *asqtad, QOP_D_Real *fatlinks[], QOP_D_Real *longlinks[], QOP_evenodd_t evenodd); /* :332-360 */
void QOP_F3_asqtad_load_L_from_raw(QOP_F3_FermionLinksAsqtad *asqtad, QOP_F_Real *fatlinks[], QOP_F_Real *longlinks[], QOP_evenodd_t evenodd);
void QOP_D3_asqtad_load_L_from_G(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad, QOP_asqtad_coeffs_t *coeffs, QOP_D3_GaugeField *gauge); /* :857-875 */
void QOP_F3_asqtad_load_L_from_G(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad, QOP_asqtad_coeffs_t *coeffs, QOP_F3_GaugeField *gauge);
void QOP_D3_asqtad_load_L_from_G2(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad, QOP_asqtad_coeffs_t *coeffs, QOP_D3_GaugeField *gFat, QOP_D3_GaugeField *gLong); /* :878-896 */
void QOP_F3_asqtad_load_L_from_G2(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad, QOP_asqtad_coeffs_t *coeffs, QOP_F3_GaugeField *gFat, QOP_F3_GaugeField *gLong);
void QOP_D3_asqtad_destroy_L(QOP_D3_FermionLinksAsqtad *field); /* :941-968 */
void QOP_F3_asqtad_destroy_L(QOP_F3_FermionLinksAsqtad *field);
void QOP_D3_asqtad_rephase_L(QOP_D3_FermionLinksAsqtad *fla, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign); /* :1448-1482 */
void QOP_F3_asqtad_rephase_L(QOP_F3_FermionLinksAsqtad *fla, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign);
/* ---- HISQ links (include/qop_d3.h:474-485; lib/hisq_links_p.c:62-301) --
 * fat7 smearing, U(3) projection, asqtad smearing of the projected links:
 * get_asqtad_links_from_hisq returns n_naiks asqtad link objects (owned by
 * the HISQ object) usable with every QOP_*_asqtad_* function. */
QOP_status_t QOP_hisq_links_set_opts(QOP_opt_t opts[], int nopts); /* lib/hisq_links.c:31-77 */
QOP_D3_FermionLinksHisq *QOP_D3_hisq_create_L_from_G(QOP_info_t *info, QOP_hisq_coeffs_t *coeffs,
                                                     QOP_D3_GaugeField *gauge); /* :62-256 */
QOP_F3_FermionLinksHisq *QOP_F3_hisq_create_L_from_G(QOP_info_t *info, QOP_hisq_coeffs_t *coeffs,
                                                     QOP_F3_GaugeField *gauge);
void QOP_D3_hisq_destroy_L(QOP_D3_FermionLinksHisq *field); /* :258-289 */
void QOP_F3_hisq_destroy_L(QOP_F3_FermionLinksHisq *field);
QOP_D3_FermionLinksAsqtad **QOP_D3_get_asqtad_links_from_hisq(QOP_D3_FermionLinksHisq *hl); /* :291-295 */
QOP_F3_FermionLinksAsqtad **QOP_F3_get_asqtad_links_from_hisq(QOP_F3_FermionLinksHisq *hl);
QOP_D3_FermionLinksAsqtad *QOP_D3_get_asqtad_deps_links_from_hisq(QOP_D3_FermionLinksHisq *hl); /* :297-301 */
QOP_F3_FermionLinksAsqtad *QOP_F3_get_asqtad_deps_links_from_hisq(QOP_F3_FermionLinksHisq *hl);

/* precision copies (asqtad_dslash_p.c:784-790, 832-838) */
QOP_F3_FermionLinksAsqtad *QOP_FD3_asqtad_create_L_from_L(QOP_D3_FermionLinksAsqtad *fla_src);
QOP_D3_FermionLinksAsqtad *QOP_DF3_asqtad_create_L_from_L(QOP_F3_FermionLinksAsqtad *fla_src);

/* ---- the hot path (include/qop_pc.h:300-371) --------------------------- */
/* out_{eo_out} = D(mass) in_{eo_in} (asqtad_dslash_p.c:1025-1092) */
void QOP_D3_asqtad_dslash(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad, QOP_D_Real mass,
                          QOP_D3_ColorVector *out, QOP_D3_ColorVector *in, QOP_evenodd_t eo_out,
                          QOP_evenodd_t eo_in);
void QOP_F3_asqtad_dslash(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad, QOP_F_Real mass,
                          QOP_F3_ColorVector *out, QOP_F3_ColorVector *in, QOP_evenodd_t eo_out,
                          QOP_evenodd_t eo_in);
/* One-direction transport (asqtad_dslash_p.c:1140-1191, 1338-1377), used
 * for interpolating operators: on the sites of eo_out,
 *   fb > 0: out += wtfat F(x)/2 in(x+dir) + wtlong L(x)/2 in(x+3 dir)
 *   fb < 0: out += s [wtfat F(x-dir)^+/2 in(x-dir) + wtlong L(x-3dir)^+/2 in(x-3dir)]
 * with s = -1 for the double-stored styles (1, 3; the default) and s = +1
 * for the single-stored ones (0, 2), as the reference's two code paths do.
 * out is accumulated into, not overwritten.  Without Naik links only the fat
 * term is applied. */
void QOP_D3_asqtad_dslash_dir(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad, int dir, int fb,
                              double wtfat, double wtlong, QOP_D3_ColorVector *out,
                              QOP_D3_ColorVector *in, QOP_evenodd_t eo_out);
void QOP_F3_asqtad_dslash_dir(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad, int dir, int fb,
                              double wtfat, double wtlong, QOP_F3_ColorVector *out,
                              QOP_F3_ColorVector *in, QOP_evenodd_t eo_out);
/* out = in/mass on eo (asqtad_dslash_p.c:1037-1104) */
void QOP_D3_asqtad_diaginv(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad, QOP_D_Real mass,
                           QOP_D3_ColorVector *out, QOP_D3_ColorVector *in, QOP_evenodd_t eo);
void QOP_F3_asqtad_diaginv(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad, QOP_F_Real mass,
                           QOP_F3_ColorVector *out, QOP_F3_ColorVector *in, QOP_evenodd_t eo);
/* asqtad_invert_p.c:134-144, 649-851 */
void QOP_D3_asqtad_invert(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad,
                          QOP_invert_arg_t *inv_arg, QOP_resid_arg_t *res_arg, QOP_D_Real mass,
                          QOP_D3_ColorVector *out_pt, QOP_D3_ColorVector *in_pt);
void QOP_F3_asqtad_invert(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad,
                          QOP_invert_arg_t *inv_arg, QOP_resid_arg_t *res_arg, QOP_F_Real mass,
                          QOP_F3_ColorVector *out_pt, QOP_F3_ColorVector *in_pt);
/* asqtad_invert_p.c:853-1059 */
void QOP_D3_asqtad_invert_multi(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *asqtad,
                                QOP_invert_arg_t *inv_arg, QOP_resid_arg_t **res_arg[],
                                QOP_D_Real *masses[], int nmass[], QOP_D3_ColorVector **out_pt[],
                                QOP_D3_ColorVector *in_pt[], int nsrc);
void QOP_F3_asqtad_invert_multi(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *asqtad,
                                QOP_invert_arg_t *inv_arg, QOP_resid_arg_t **res_arg[],
                                QOP_F_Real *masses[], int nmass[], QOP_F3_ColorVector **out_pt[],
                                QOP_F3_ColorVector *in_pt[], int nsrc);

/* ==== extensions (no counterpart in the reference) ===================== */

/* Seeded inputs in the raw layout (host).  Gauge: U = reunit(exp(i eps H)),
 * H traceless Hermitian Gaussian; vector: unit-variance complex Gaussian on
 * the sites of `evenodd`, zero elsewhere. */
void QOPX_gen_gauge(const int L[4], unsigned long long seed, double eps, double *U[4]);
void QOPX_gen_gaussian_V(const int L[4], unsigned long long seed, int evenodd, double *v);

/* Batched mixed-precision solves: out[i] = m_i A_i^-1 in[i] on the EVEN or
 * ODD subset (inv_arg->evenodd), each to FP64 true relative residual
 * sqrt(res[i]->rsqmin).  Up to `capacity` systems iterate together and a
 * finished slot takes the next system (continuous batching).  prec: sloppy
 * precision 32 or 16 (FP16 vectors with per-site norms, FP16 links with
 * per-link scales).  latency[i] (optional): seconds from the call's start to
 * system i's completion.  Returns the number of systems that converged, or
 * -1 with info->status = QOP_MEM_ERROR when the batch workspace does not fit
 * in device memory (nothing is kept allocated then). */
int QOPX_D3_asqtad_invert_batch(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *fla,
                                QOP_invert_arg_t *inv_arg, QOP_resid_arg_t *res[], double masses[],
                                QOP_D3_ColorVector *out[], QOP_D3_ColorVector *in[], int n,
                                int capacity, int prec, double *latency);

/* Multi-mass solves of nsrc sources over one mass list (nmass <= 16):
 * out[i*nmass+j] = m_j A_j^-1 in[i] on the EVEN or ODD subset, each to FP64
 * true relative residual sqrt(res[i*nmass+j]->rsqmin).  An FP32 multi-shift
 * CG with FP64 reliable updates, batched over the sources, reaches relative
 * residual ms_tol (0 selects the smallest requested tolerance); every
 * (source, mass) whose FP64 true residual is still above its tolerance is
 * then refined by the reliable-update CG in sloppy precision prec (16 or
 * 32), `capacity` systems at a time.
 * ms_iters[i] (optional): multi-shift iterations of source i.  Returns the
 * number of systems that converged, or -1 with info->status = QOP_MEM_ERROR
 * when the device memory is exhausted. */
int QOPX_D3_asqtad_invert_multi_batch(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *fla,
                                      QOP_invert_arg_t *inv_arg, QOP_resid_arg_t *res[],
                                      double masses[], int nmass, QOP_D3_ColorVector *out[],
                                      QOP_D3_ColorVector *in[], int nsrc, int capacity, int prec,
                                      double ms_tol, int *ms_iters);

/* The hop out = D_{eo_out, 1-eo_out} in as the sloppy solvers apply it, in
 * precision prec (64, 32, or 16: FP16 vectors with a per-site float norm and
 * FP16 links with a per-link float scale), on n (1..64) right-hand sides in
 * one multi-RHS launch of the autotuned shape; in[i] is rounded to the stored
 * precision on the source parity, out[i] receives the result on eo_out in
 * FP64.  Returns 0, or -1 when the device memory is exhausted. */
int QOPX_D3_asqtad_hop(QOP_D3_FermionLinksAsqtad *fla, int prec, int n, QOP_D3_ColorVector *out[],
                       QOP_D3_ColorVector *in[], QOP_evenodd_t eo_out);

/* QOPX_D3_asqtad_hop without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * stream capture, serving loops that must not reach the allocator).  Same
 * kernels and bit-identical results; the stored-precision source and result
 * batches live in a 64 MiB device array held by the library for the whole
 * process on every device it runs on, so 2 x n vectors of Vh sites in the
 * stored precision (48 B per site in FP64, 24 B in FP32, 16 B in FP16) must fit
 * in 64 MiB.  The links in precision prec must exist already (a
 * QOPX_D3_asqtad_hop call in that precision builds them); the launch shape
 * is the tuned one when that call tuned it, a plain one otherwise.  Calls on
 * one device share the array: calls that overlap in time on one device race.
 * Returns 0, or -2 when the vectors do not fit or the links in that precision
 * are not built (nothing is computed then). */
int QOPX_D3_asqtad_hop_noalloc(QOP_D3_FermionLinksAsqtad *fla, int prec, int n,
                               QOP_D3_ColorVector *out[], QOP_D3_ColorVector *in[],
                               QOP_evenodd_t eo_out);

/* Mean ms per launch of the even->odd hop kernel (precision 64/32/16, nrhs
 * right-hand sides, autotuned shape (rhs per thread, block, XCD remap,
 * non-temporal links, RHS-innermost order, x0 tile, LDS staging, plain loads)
 * returned in shape[8]), from HIP events.  red = 0: plain hop; 1: with the
 * fused |out|^2 reduction (first hop of a batched CG iteration); 2: with the
 * fused CG update r -= a (m^2 p - D t), x += a p, |r|^2 (second hop;
 * precision 32/16).  -1 if the operands do not fit in device memory. */
double QOPX_D3_asqtad_hop_time(QOP_D3_FermionLinksAsqtad *fla, int prec, int nrhs, int nrep,
                               int red, int *shape);

/* Force the multi-RHS hop launch shape {rhs per thread, block, XCD remap,
 * non-temporal links, RHS-innermost order, x0 tile, LDS staging, plain loads
 * (1: FP32/FP16 through the one-RHS-per-thread plain-load kernel that
 * lattices beyond the 2^31-byte buffer range take; 0: buffer loads)} of
 * solvers built from now on, where it is valid for their lattice and
 * precision (tests of launch-shape independence); null restores autotuning. */
void QOPX_hop_shape_override(const int *shape);

/* Mean ms per launch of a plain device copy of `bytes` (multiple of 16):
 * the measured copy bandwidth and the calibration kernel for HBM counters.
 * -1 if the buffers do not fit (as QOPX_D3_asqtad_hop_time). */
double QOPX_copy_time(unsigned long long bytes, int nrep);

/* ---- asqtad fermion force (include/qop_pc.h:427-465; lib/asqtad_force_p.c) --
 * force[mu] += -2 TA(U_mu deriv_mu^+) with deriv the derivative of
 * sum_i eps[i] x_i x_i^+ through the fat7 + Lepage + Naik smearing
 * (lib/asqtad_force_fnmat_p.c:474-546); in_pt[i] are full-lattice vectors,
 * x_i on even sites and D_oe x_i on odd sites as the caller prepares them.
 * gauge is the rephased gauge field the links were smeared from.  Every nsrc
 * takes the fnmat form on the device (the reference's ASVEC form for
 * nsrc < fnmat_src_min evaluates the same derivative).  FP64 arithmetic for
 * both precisions.  info->status = QOP_MEM_ERROR if the workspace does not
 * fit in device memory (force unchanged then). */
QOP_status_t QOP_asqtad_force_set_opts(QOP_opt_t opts[], int nopts); /* lib/asqtad_force.c:22-43 */
void QOP_D3_asqtad_force(QOP_info_t *info, QOP_D3_GaugeField *gauge, QOP_D3_Force *force,
                         QOP_asqtad_coeffs_t *coeffs, QOP_D_Real eps,
                         QOP_D3_ColorVector *in_pt); /* lib/asqtad_force_p.c:55-63 */
void QOP_F3_asqtad_force(QOP_info_t *info, QOP_F3_GaugeField *gauge, QOP_F3_Force *force,
                         QOP_asqtad_coeffs_t *coeffs, QOP_F_Real eps, QOP_F3_ColorVector *in_pt);
void QOP_D3_asqtad_force_multi(QOP_info_t *info, QOP_D3_GaugeField *gauge, QOP_D3_Force *force,
                               QOP_asqtad_coeffs_t *coef, QOP_D_Real eps[],
                               QOP_D3_ColorVector *in_pt[], int nsrc); /* lib/asqtad_force_p.c:41-53 */
void QOP_F3_asqtad_force_multi(QOP_info_t *info, QOP_F3_GaugeField *gauge, QOP_F3_Force *force,
                               QOP_asqtad_coeffs_t *coef, QOP_F_Real eps[],
                               QOP_F3_ColorVector *in_pt[], int nsrc);

/* ---- HISQ fermion force (include/qop_pc.h:485-510; lib/hisq_force_p.c) ----
 * force[mu] += TA(U_mu deriv_mu^+), deriv the derivative of
 * sum_i eps[i] x_i x_i^+ through the level-2 smearing of W (per
 * epsilon-Naik group: in_pt grouped by n_orders_naik[0..n_naiks-1]), the
 * U(3) projection W = V (V^+V)^-1/2 and the fat7 level on U
 * (lib/hisq_force_fnmat_p.c:7-231).  Needs the auxiliary links kept by
 * QOP_hisq_create_L_from_G (option want_aux, default 1) and
 * umethod = QOP_UNITARIZE_RATIONAL; QOP_FAIL otherwise.  The number of links
 * whose eigenvalues the force filter raised is added to
 * QOP_info_hisq_force_filter_counter(info).  Options (lib/hisq_force.c:29-52):
 * fnmat_src_min, veclength (kept; every nsrc takes the fnmat form),
 * force_filter (default 5e-5). */
#define QOP_info_hisq_force_filter_counter(info) ((info)->count2)
QOP_status_t QOP_hisq_force_set_opts(QOP_opt_t opts[], int nopts); /* lib/hisq_force.c:29-52 */
void QOP_D3_hisq_force_multi(QOP_info_t *info, QOP_D3_FermionLinksHisq *flh, QOP_D3_Force *force,
                             QOP_hisq_coeffs_t *coef, QOP_D_Real eps[], QOP_D3_ColorVector *in_pt[],
                             int *n_orders_naik); /* lib/hisq_force_p.c:63-83 */
void QOP_F3_hisq_force_multi(QOP_info_t *info, QOP_F3_FermionLinksHisq *flh, QOP_F3_Force *force,
                             QOP_hisq_coeffs_t *coef, QOP_F_Real eps[], QOP_F3_ColorVector *in_pt[],
                             int *n_orders_naik);

/* ---- request server: continuous batching over resident configurations ----
 * A server keeps fat/long links of any number of gauge configurations, of
 * any lattice sizes, resident on the device (QOP_init is not needed and its
 * lattice is not used) and solves EVEN/ODD requests out = m A^-1 src to FP64
 * true relative residual `tol`.  A host worker thread admits requests as
 * slots and HBM free up into per-(configuration, parity) batches that are
 * already iterating, so requests that arrive while others run join them at
 * the next iteration chunk; batches of different configurations run
 * concurrently on separate streams.  A request's result does not depend on
 * what it was batched with.
 *   capacity  slots per batch (1..256)
 *   prec      16 or 32: sloppy precision of the reliable-update CG;
 *             64: the FP64 reference algorithm, one request at a time
 *   max_bytes HBM budget for links and batches (0: the device's free memory
 *             less a reserve); a batch is created only when it fits, idle
 *             batches of other configurations are evicted to make room.
 * Returns null when no device is usable or an argument is invalid. */
typedef struct QOPX_Server_struct QOPX_Server;
QOPX_Server *QOPX_serve_create(int capacity, int prec, unsigned long long max_bytes);
/* Add a configuration from a raw FP64 gauge field (host): rephased with r0,
 * bc, sign (as QOP_D3_rephase_G; null pointers skip) and smeared with coeffs
 * on the device.  Returns the configuration index, or -1 if over budget or
 * out of device memory. */
int QOPX_serve_add_gauge(QOPX_Server *s, const int L[4], QOP_D_Real *U[4],
                         QOP_asqtad_coeffs_t *coeffs, int *r0, QOP_bc_t *bc,
                         QOP_staggered_sign_t *sign);
/* Add a configuration from raw FP64 fat and long links (host, as
 * QOP_D3_asqtad_create_L_from_raw; lng may be null). */
int QOPX_serve_add_links(QOPX_Server *s, const int L[4], QOP_D_Real *fat[4], QOP_D_Real *lng[4]);
/* Queue a solve on subset evenodd (QOP_EVEN or QOP_ODD) of configuration
 * `config`.  src and out are raw FP64 host vectors [V][3][2] that must stay
 * valid until the request is waited for; only the sites of evenodd of out are
 * written.  max_iter <= 0 selects 20000.  Returns a request id, or -1. */
long QOPX_serve_submit(QOPX_Server *s, int config, QOP_evenodd_t evenodd, double mass, double tol,
                       int max_iter, const QOP_D_Real *src, QOP_D_Real *out);
/* Block until the request is done; fill res (final_rsq = FP64 true
 * |r|^2/|src|^2, final_iter, final_restart = reliable updates) and latency
 * (seconds from submit to completion).  Returns 0 if it converged, 1 if it
 * stopped at max_iter, -1 for an unknown or failed request.  A request is
 * waited for once. */
int QOPX_serve_wait(QOPX_Server *s, long request, QOP_resid_arg_t *res, double *latency);
/* 1 if the request is done, 0 if not, -1 if unknown. */
int QOPX_serve_test(QOPX_Server *s, long request);
/* completed, queued, running, batches, bytes used, byte budget, chunks
 * launched, failed. */
void QOPX_serve_stats(QOPX_Server *s, double stats[8]);
/* Finish every admitted request, then free the server.  Requests that can
 * never be admitted fail. */
void QOPX_serve_destroy(QOPX_Server *s);

/* Library status: 1 when a HIP device is usable. */
int QOPX_device_ok(void);
/* Page-locked host memory for raw arrays (vectors, links, gauge fields):
 * create/extract/convert calls and request-server transfers on such buffers
 * run at the full host-link rate (direct DMA) instead of the staged pageable
 * rate.  NULL if the allocation fails; free with QOPX_host_free. */
void *QOPX_host_alloc(unsigned long long bytes);
void QOPX_host_free(void *p);
/* Last error message (empty when none). */
const char *QOPX_last_error(void);

/* ==== precision-generic names ==========================================
 * As in the reference (include/qop.h:8-48 selects the precision, default
 * single; include/qop_d3_generic.h / qop_f3_generic.h map the names), a
 * program built with QOP_Precision 'D' or 'F' (or QOP_PrecisionInt 2 / 1)
 * may call the generic names below.  Only Nc = 3 exists here. */
#ifndef QOP_MI355X_NO_GENERIC
#ifndef QOP_Precision
#  ifndef QOP_PrecisionInt
#    define QOP_PrecisionInt 1
#  endif
#  if QOP_PrecisionInt == 1
#    define QOP_Precision 'F'
#  elif QOP_PrecisionInt == 2
#    define QOP_Precision 'D'
#  else
#    error "bad QOP_PrecisionInt"
#  endif
#endif
#if QOP_Precision == 'D'
#  define QOP_Real QOP_D_Real
#  define QOP_ColorVector QOP_D3_ColorVector
#  define QOP_GaugeField QOP_D3_GaugeField
#  define QOP_FermionLinksAsqtad QOP_D3_FermionLinksAsqtad
#  define QOP_create_V_from_raw QOP_D3_create_V_from_raw
#  define QOP_extract_V_to_raw QOP_D3_extract_V_to_raw
#  define QOP_destroy_V QOP_D3_destroy_V
#  define QOP_convert_V_from_raw QOP_D3_convert_V_from_raw
#  define QOP_convert_V_to_raw QOP_D3_convert_V_to_raw
#  define QOP_create_G_from_raw QOP_D3_create_G_from_raw
#  define QOP_extract_G_to_raw QOP_D3_extract_G_to_raw
#  define QOP_destroy_G QOP_D3_destroy_G
#  define QOP_rephase_G QOP_D3_rephase_G
#  define QOP_convert_G_from_raw QOP_D3_convert_G_from_raw
#  define QOP_convert_G_to_raw QOP_D3_convert_G_to_raw
#  define QOP_convert_F_from_raw QOP_D3_convert_F_from_raw
#  define QOP_convert_F_to_raw QOP_D3_convert_F_to_raw
#  define QOP_asqtad_create_L_from_raw QOP_D3_asqtad_create_L_from_raw
#  define QOP_asqtad_create_L_from_G QOP_D3_asqtad_create_L_from_G
#  define QOP_asqtad_create_L_from_G2 QOP_D3_asqtad_create_L_from_G2
#  define QOP_asqtad_extract_L_to_raw QOP_D3_asqtad_extract_L_to_raw
#  define QOP_asqtad_convert_L_from_raw QOP_D3_asqtad_convert_L_from_raw
#  define QOP_asqtad_convert_L_to_raw QOP_D3_asqtad_convert_L_to_raw
#  define QOP_asqtad_load_L_from_G2 QOP_D3_asqtad_load_L_from_G2
#  define QOP_asqtad_dslash_dir QOP_D3_asqtad_dslash_dir
#  define QOP_asqtad_load_L_from_raw QOP_D3_asqtad_load_L_from_raw
#  define QOP_asqtad_load_L_from_G QOP_D3_asqtad_load_L_from_G
#  define QOP_asqtad_destroy_L QOP_D3_asqtad_destroy_L
#  define QOP_asqtad_rephase_L QOP_D3_asqtad_rephase_L
#  define QOP_asqtad_dslash QOP_D3_asqtad_dslash
#  define QOP_asqtad_diaginv QOP_D3_asqtad_diaginv
#  define QOP_asqtad_invert QOP_D3_asqtad_invert
#  define QOP_asqtad_invert_multi QOP_D3_asqtad_invert_multi
#  define QOP_FermionLinksHisq QOP_D3_FermionLinksHisq
#  define QOP_hisq_create_L_from_G QOP_D3_hisq_create_L_from_G
#  define QOP_hisq_destroy_L QOP_D3_hisq_destroy_L
#  define QOP_get_asqtad_links_from_hisq QOP_D3_get_asqtad_links_from_hisq
#  define QOP_get_asqtad_deps_links_from_hisq QOP_D3_get_asqtad_deps_links_from_hisq
#  define QOP_Force QOP_D3_Force
#  define QOP_create_F_from_raw QOP_D3_create_F_from_raw
#  define QOP_extract_F_to_raw QOP_D3_extract_F_to_raw
#  define QOP_destroy_F QOP_D3_destroy_F
#  define QOP_asqtad_force QOP_D3_asqtad_force
#  define QOP_asqtad_force_multi QOP_D3_asqtad_force_multi
#  define QOP_hisq_force_multi QOP_D3_hisq_force_multi
#elif QOP_Precision == 'F'
#  define QOP_Real QOP_F_Real
#  define QOP_ColorVector QOP_F3_ColorVector
#  define QOP_GaugeField QOP_F3_GaugeField
#  define QOP_FermionLinksAsqtad QOP_F3_FermionLinksAsqtad
#  define QOP_create_V_from_raw QOP_F3_create_V_from_raw
#  define QOP_extract_V_to_raw QOP_F3_extract_V_to_raw
#  define QOP_destroy_V QOP_F3_destroy_V
#  define QOP_convert_V_from_raw QOP_F3_convert_V_from_raw
#  define QOP_convert_V_to_raw QOP_F3_convert_V_to_raw
#  define QOP_create_G_from_raw QOP_F3_create_G_from_raw
#  define QOP_extract_G_to_raw QOP_F3_extract_G_to_raw
#  define QOP_destroy_G QOP_F3_destroy_G
#  define QOP_rephase_G QOP_F3_rephase_G
#  define QOP_convert_G_from_raw QOP_F3_convert_G_from_raw
#  define QOP_convert_G_to_raw QOP_F3_convert_G_to_raw
#  define QOP_convert_F_from_raw QOP_F3_convert_F_from_raw
#  define QOP_convert_F_to_raw QOP_F3_convert_F_to_raw
#  define QOP_asqtad_create_L_from_raw QOP_F3_asqtad_create_L_from_raw
#  define QOP_asqtad_create_L_from_G QOP_F3_asqtad_create_L_from_G
#  define QOP_asqtad_extract_L_to_raw QOP_F3_asqtad_extract_L_to_raw
#  define QOP_asqtad_convert_L_from_raw QOP_F3_asqtad_convert_L_from_raw
#  define QOP_asqtad_convert_L_to_raw QOP_F3_asqtad_convert_L_to_raw
#  define QOP_asqtad_load_L_from_G2 QOP_F3_asqtad_load_L_from_G2
#  define QOP_asqtad_dslash_dir QOP_F3_asqtad_dslash_dir
#  define QOP_asqtad_create_L_from_G2 QOP_F3_asqtad_create_L_from_G2
#  define QOP_asqtad_load_L_from_raw QOP_F3_asqtad_load_L_from_raw
#  define QOP_asqtad_load_L_from_G QOP_F3_asqtad_load_L_from_G
#  define QOP_asqtad_rephase_L QOP_F3_asqtad_rephase_L
#  define QOP_asqtad_destroy_L QOP_F3_asqtad_destroy_L
#  define QOP_asqtad_dslash QOP_F3_asqtad_dslash
#  define QOP_asqtad_diaginv QOP_F3_asqtad_diaginv
#  define QOP_asqtad_invert QOP_F3_asqtad_invert
#  define QOP_asqtad_invert_multi QOP_F3_asqtad_invert_multi
#  define QOP_FermionLinksHisq QOP_F3_FermionLinksHisq
#  define QOP_hisq_create_L_from_G QOP_F3_hisq_create_L_from_G
#  define QOP_hisq_destroy_L QOP_F3_hisq_destroy_L
#  define QOP_get_asqtad_links_from_hisq QOP_F3_get_asqtad_links_from_hisq
#  define QOP_get_asqtad_deps_links_from_hisq QOP_F3_get_asqtad_deps_links_from_hisq
#  define QOP_Force QOP_F3_Force
#  define QOP_create_F_from_raw QOP_F3_create_F_from_raw
#  define QOP_extract_F_to_raw QOP_F3_extract_F_to_raw
#  define QOP_destroy_F QOP_F3_destroy_F
#  define QOP_asqtad_force QOP_F3_asqtad_force
#  define QOP_asqtad_force_multi QOP_F3_asqtad_force_multi
#  define QOP_hisq_force_multi QOP_F3_hisq_force_multi
#else
#  error "bad QOP_Precision"
#endif
#endif /* QOP_MI355X_NO_GENERIC */

#ifdef __cplusplus
}
#endif
#endif /* QOP_MI355X_H */
