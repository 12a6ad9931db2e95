// The QOP C ABI (asqtad subset, Nc = 3) on top of the device runtime.
// Each entry point follows the reference function cited beside it.
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "handles.h"
#include "solver.h"

using namespace qop;

namespace {

struct Global {
  int inited = 0;
  int verbosity = 0;
  int proflevel = 0;
  Lattice *lat = nullptr;
} G;

std::string g_last_error;

void check_init(const char *fn) {
  if (!G.inited) {
    fprintf(stderr, "Error: QOP not initialized in function %s\n", fn);
    exit(1);  // CHECK_INIT, include/qop_internal.h:67-71
  }
}
#define CHECK_INIT check_init(__func__)

void eo_range(const Lattice *lat, int eo, long &b, long &e) {
  b = eo == QOP_ODD ? lat->Vh : 0;
  e = eo == QOP_EVEN ? lat->Vh : lat->V;
}

// Upload the eo sites of a host raw array (T = double/float) into a device
// vector.  Sites outside eo keep their previous contents.
template <typename T>
void upload_vec(DevVec &v, const T *raw, int eo) {
  Lattice *lat = v.lat;
  long b, e;
  eo_range(lat, eo, b, e);
  DevBuf<T> buf((size_t)lat->V * 6 * sizeof(T));
  T *d = buf.p;
  QOP_CHECK_HIP(hipMemcpyAsync(d + 6 * b, raw + 6 * b, (e - b) * 6 * sizeof(T), hipMemcpyHostToDevice, stream()));
  if constexpr (sizeof(T) == 8) launch_raw_to_vec(lat, (const double *)d, v, eo, stream());
  else launch_rawf_to_vec(lat, (const float *)d, v, eo, stream());
  sync();
}

template <typename T>
void download_vec(T *raw, const DevVec &v, int eo) {
  Lattice *lat = v.lat;
  long b, e;
  eo_range(lat, eo, b, e);
  DevBuf<T> buf((size_t)lat->V * 6 * sizeof(T));
  T *d = buf.p;
  if constexpr (sizeof(T) == 8) launch_vec_to_raw(lat, v, (double *)d, eo, stream());
  else launch_vec_to_rawf(lat, v, (float *)d, eo, stream());
  QOP_CHECK_HIP(hipMemcpyAsync(raw + 6 * b, d + 6 * b, (e - b) * 6 * sizeof(T), hipMemcpyDeviceToHost, stream()));
  sync();
}

// Host raw matrices (T) -> device FP64 raw matrices.
template <typename T>
double *upload_mat(Lattice *lat, const T *raw, int eo) {
  long b, e;
  eo_range(lat, eo, b, e);
  size_t n = (size_t)lat->V * 18;
  double *d;
  dev_alloc(&d, n * sizeof(double));
  QOP_CHECK_HIP(hipMemsetAsync(d, 0, n * sizeof(double), stream()));
  if constexpr (sizeof(T) == 8) {
    QOP_CHECK_HIP(hipMemcpyAsync(d + 18 * b, raw + 18 * b, (e - b) * 18 * sizeof(double), hipMemcpyHostToDevice, stream()));
    sync();
  } else {
    std::vector<double> tmp((e - b) * 18);
    for (long i = 0; i < (e - b) * 18; i++) tmp[i] = raw[18 * b + i];
    QOP_CHECK_HIP(hipMemcpyAsync(d + 18 * b, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, stream()));
    sync();
  }
  return d;
}

template <typename T>
void download_mat(T *raw, const double *d, Lattice *lat, int eo) {
  long b, e;
  eo_range(lat, eo, b, e);
  std::vector<double> tmp((e - b) * 18);
  QOP_CHECK_HIP(hipMemcpy(tmp.data(), d + 18 * b, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < tmp.size(); i++) raw[18 * b + i] = (T)tmp[i];
}

template <typename T>
void build_links_from_raw(Links &L, T *fat[], T *lng[], int eo) {
  const double *f[4], *l[4];
  DevBuf<double> fd[4], ld[4];
  for (int i = 0; i < 4; i++) {
    fd[i].p = upload_mat(L.lat, fat[i], eo);
    f[i] = fd[i].p;
    if (lng) { ld[i].p = upload_mat(L.lat, lng[i], eo); l[i] = ld[i].p; }
  }
  build_links64(L, f, lng ? l : nullptr, stream());
}

void fatten(Links &L, Lattice *lat, const double *const gf[4], const double *const gl[4],
            const QOP_asqtad_coeffs_t *c, QOP_info_t *info) {
  double coef[6] = {c->one_link, c->three_staple, c->five_staple, c->seven_staple, c->lepage, c->naik};
  bool naik = c->naik != 0;
  DevBuf<double> fbuf[4], lbuf[4];
  double *fat[4], *lng[4];
  size_t n = (size_t)lat->V * 18 * sizeof(double);
  for (int i = 0; i < 4; i++) {
    dev_alloc(&fbuf[i].p, n);
    if (naik) dev_alloc(&lbuf[i].p, n);
    fat[i] = fbuf[i].p;
    lng[i] = lbuf[i].p;
  }
  const double *g[4] = {gf[0], gf[1], gf[2], gf[3]}, *gL[4] = {gl[0], gl[1], gl[2], gl[3]};
  double t0 = wall_time();
  launch_fatten(lat, g, gL, coef, fat, naik ? lng : nullptr, stream());
  double dt = wall_time() - t0;
  L.lat = lat;
  const double *fc[4] = {fat[0], fat[1], fat[2], fat[3]}, *lc[4] = {lng[0], lng[1], lng[2], lng[3]};
  build_links64(L, fc, naik ? lc : nullptr, stream());
  bool have3 = c->three_staple != 0 || c->five_staple != 0 || c->seven_staple != 0 || c->lepage != 0;
  if (info) {  // asqtad_dslash_p.c:536, 624-626
    info->final_sec = dt;
    info->final_flop = (have3 || naik) ? 61632.0 * lat->V : 0.0;
    info->status = QOP_SUCCESS;
  }
}

void rephase_dev(Lattice *lat, double *const U[4], int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign,
                 int fat_slices_only, int nslice_long, double *const Ulong[4], int sign_rel) {
  int z[4] = {0, 0, 0, 0};
  int *rr = r0 ? r0 : z;
  for (int mu = 0; mu < 4; mu++) {
    int bcdir = bc && (bc->phase[mu].re != 1. || bc->phase[mu].im != 0.);
    double br = bc ? bc->phase[mu].re : 1, bi = bc ? bc->phase[mu].im : 0;
    int bits = sign ? sign->signmask[mu] : 0;
    if (!bcdir && !bits) continue;
    launch_rephase(lat, U[mu], mu, rr, br, bi, bcdir, bits, 1, sign_rel, stream());
    if (Ulong && Ulong[mu]) launch_rephase(lat, Ulong[mu], mu, rr, br, bi, bcdir, bits, nslice_long, sign_rel, stream());
  }
  (void)fat_slices_only;
  sync();
}

}  // namespace

namespace qop {

void links_from_raw64(Links &L, double *fat[4], double *lng[4]) {
  build_links_from_raw(L, fat, lng, QOP_EVENODD);
  sync();
}

void links_from_gauge64(Links &L, double *const U_host[4], const QOP_asqtad_coeffs_t *c, int *r0,
                        QOP_bc_t *bc, QOP_staggered_sign_t *sign) {
  double *U[4];
  for (int mu = 0; mu < 4; mu++) U[mu] = upload_mat(L.lat, U_host[mu], QOP_EVENODD);
  rephase_dev(L.lat, U, r0, bc, sign, 1, 1, nullptr, 0);
  fatten(L, L.lat, U, U, c, nullptr);
  sync();
  for (int mu = 0; mu < 4; mu++) QOP_CHECK_HIP(hipFree(U[mu]));
}

}  // namespace qop

extern "C" {

// ---- general ------------------------------------------------------------
QOP_status_t QOP_init(QOP_layout_t *layout) {
  if (G.inited) {
    fprintf(stderr, "Error: QOP is initialized in function %s\n", __func__);
    exit(1);
  }
  if (!layout || layout->latdim != 4) return QOP_FAIL;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) {
    g_last_error = "no HIP device";
    fprintf(stderr, "QOP-MI355X: no HIP device available\n");
    return QOP_FAIL;
  }
  G.inited = 1;
  G.verbosity = 0;
  G.lat = lattice_get(layout->latsize);
  return QOP_SUCCESS;
}

QOP_status_t QOP_finalize(void) {
  CHECK_INIT;
  QOP_CHECK_HIP(hipDeviceSynchronize());
  G.inited = 0;
  return QOP_SUCCESS;
}

int QOP_is_initialized(void) { return G.inited; }
const char *QOP_version_str(void) { return "0.21.1-mi355x"; }
int QOP_version_int(void) { return (0 * 1000 + 21) * 1000 + 1; }
int QOP_verbose(int level) {
  int old = G.verbosity;
  G.verbosity = level;
  options().verbose = level >= QOP_VERB_MED;
  return old;
}
int QOP_profcontrol(int level) {
  int old = G.proflevel;
  G.proflevel = level;
  return old;
}
int QOP_node_number_raw(int coords[]) {
  CHECK_INIT;
  (void)coords;
  return 0;
}
static int raw_index(const int x[4]) {
  const int *L = G.lat->L;
  long l = ((long)(x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3];
  int p = (x[0] + x[1] + x[2] + x[3]) & 1;
  return (int)(p ? (l + G.lat->V) / 2 : l / 2);
}
int QOP_node_index_raw_V(int coords[], QOP_evenodd_t) { CHECK_INIT; return raw_index(coords); }
int QOP_node_index_raw_G(int coords[], QOP_evenodd_t) { CHECK_INIT; return raw_index(coords); }
int QOP_sites_on_node_raw_V(QOP_evenodd_t) { CHECK_INIT; return (int)G.lat->V; }
int QOP_sites_on_node_raw_G(QOP_evenodd_t) { CHECK_INIT; return (int)G.lat->V; }

// lib/asqtad_invert.c:68-112 plus "qmx_*" tags of this library
QOP_status_t QOP_asqtad_invert_set_opts(QOP_opt_t opts[], int nopts) {
  CHECK_INIT;
  Options &o = options();
  int st = o.style, ns = o.nsvec, nm = o.nvec;
  auto get = [&](const char *tag, auto &var) {
    for (int i = 0; i < nopts; i++)
      if (!strcmp(opts[i].tag, tag)) var = (std::remove_reference_t<decltype(var)>)opts[i].value;
  };
  auto valid_nvec = [](int nv, int s) { return nv == 1 || nv == 2 || nv == 4 || nv == 8 || (s == 1 && nv == 16); };
  get("st", st);
  if (st < 0 || st > 3) return QOP_FAIL;
  if (!valid_nvec(ns, st)) ns = ns < 1 ? 1 : 8;
  if (!valid_nvec(nm, st)) nm = nm < 1 ? 1 : 8;
  get("ns", ns);
  get("nm", nm);
  if (!valid_nvec(ns, st) || !valid_nvec(nm, st)) return QOP_FAIL;
  int cg = o.cgtype, nev = o.eigcg_nev, m = o.eigcg_m, numax = o.eigcg_numax;
  get("cg", cg);
  get("eigcg_nev", nev);
  get("eigcg_m", m);
  get("eigcg_l", numax);
  int solver = o.solver, prec = o.inner_prec == P32 ? 32 : o.inner_prec == P16 ? 16 : 64;
  double delta = o.delta;
  int chunk = o.chunk, graph = o.use_graph;
  get("qmx_solver", solver);
  get("qmx_prec", prec);
  get("qmx_delta", delta);
  get("qmx_chunk", chunk);
  get("qmx_graph", graph);
  int msp = o.ms_prec == P16 ? 16 : 32, msl = o.ms_link16 ? 16 : 32;
  get("qmx_ms_prec", msp);
  get("qmx_ms_links", msl);
  double join = o.ru_join;
  get("qmx_ru_join", join);
  if (!(join >= 0 && join < 1)) return QOP_FAIL;
  // eigCG sizes (not checked by the reference, lib/asqtad_invert.c:94-107):
  // the restart keeps 2 nev + 1 Lanczos vectors of m, and the space holds m
  if (cg == 1 && (nev < 1 || m < 2 * nev + 2 || numax < m)) return QOP_FAIL;
  if (msp != 16 && msp != 32) return QOP_FAIL;
  if (msl != 16 && msl != 32) return QOP_FAIL;
  if (solver < 0 || solver > 1) return QOP_FAIL;
  if (prec != 64 && prec != 32 && prec != 16) return QOP_FAIL;
  if (!(delta > 0 && delta < 1) || chunk < 1 || chunk > 1024) return QOP_FAIL;
  o.style = st; o.nsvec = ns; o.nvec = nm; o.cgtype = cg;
  o.eigcg_nev = nev; o.eigcg_m = m; o.eigcg_numax = numax;
  o.solver = solver;
  o.inner_prec = prec == 64 ? P64 : prec == 32 ? P32 : P16;
  o.delta = delta;
  o.chunk = chunk;
  o.use_graph = graph;
  o.ms_prec = msp == 16 ? P16 : P32;
  o.ms_link16 = msl == 16;
  o.ru_join = join;
  return QOP_SUCCESS;
}

// ---- vectors --------------------------------------------------------------
QOP_D3_ColorVector *QOP_D3_create_V_from_raw(QOP_D_Real *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  auto *v = new QOP_D3_ColorVector;
  if (!guard_oom(__func__, nullptr, [&] {
        devvec_alloc(v->v, G.lat, P64);
        upload_vec(v->v, src, eo);
      })) {
    devvec_free(v->v);
    delete v;
    return nullptr;
  }
  return v;
}
QOP_F3_ColorVector *QOP_F3_create_V_from_raw(QOP_F_Real *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  auto *v = new QOP_F3_ColorVector;
  if (!guard_oom(__func__, nullptr, [&] {
        devvec_alloc(v->v, G.lat, P32);
        upload_vec(v->v, src, eo);
      })) {
    devvec_free(v->v);
    delete v;
    return nullptr;
  }
  return v;
}
void QOP_D3_extract_V_to_raw(QOP_D_Real *dest, QOP_D3_ColorVector *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  guard_oom(__func__, nullptr, [&] { download_vec(dest, src->v, eo); });
}
void QOP_F3_extract_V_to_raw(QOP_F_Real *dest, QOP_F3_ColorVector *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  guard_oom(__func__, nullptr, [&] { download_vec(dest, src->v, eo); });
}
void QOP_D3_destroy_V(QOP_D3_ColorVector *f) {
  if (!f) return;
  sync();
  devvec_free(f->v);
  delete f;
}
void QOP_F3_destroy_V(QOP_F3_ColorVector *f) {
  if (!f) return;
  sync();
  devvec_free(f->v);
  delete f;
}
QOP_D3_ColorVector *QOP_D3_convert_V_from_raw(QOP_D_Real *src, QOP_evenodd_t eo) {
  QOP_D3_ColorVector *v = QOP_D3_create_V_from_raw(src, eo);
  if (v) v->raw = src;
  return v;
}
QOP_F3_ColorVector *QOP_F3_convert_V_from_raw(QOP_F_Real *src, QOP_evenodd_t eo) {
  QOP_F3_ColorVector *v = QOP_F3_create_V_from_raw(src, eo);
  if (v) v->raw = src;
  return v;
}
// common_p.c:225-240; the raw array is sized for all sites (the reference
// allocates sites_on_node reals only, a bug not reproduced here).
QOP_D_Real *QOP_D3_convert_V_to_raw(QOP_D3_ColorVector *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  double *r = src->raw;
  const bool own = !r;
  if (own) r = (double *)calloc((size_t)src->v.lat->V * 6, sizeof(double));
  if (!guard_oom(__func__, nullptr, [&] { download_vec(r, src->v, eo); })) {
    if (own) free(r);
    return nullptr;  // src stays valid
  }
  QOP_D3_destroy_V(src);
  return r;
}
QOP_F_Real *QOP_F3_convert_V_to_raw(QOP_F3_ColorVector *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  float *r = src->raw;
  const bool own = !r;
  if (own) r = (float *)calloc((size_t)src->v.lat->V * 6, sizeof(float));
  if (!guard_oom(__func__, nullptr, [&] { download_vec(r, src->v, eo); })) {
    if (own) free(r);
    return nullptr;  // src stays valid
  }
  QOP_F3_destroy_V(src);
  return r;
}

// ---- gauge fields -----------------------------------------------------------
QOP_D3_GaugeField *QOP_D3_create_G_from_raw(QOP_D_Real *links[], QOP_evenodd_t eo) {
  CHECK_INIT;
  auto *g = new QOP_D3_GaugeField;
  g->lat = G.lat;
  if (!guard_oom(__func__, nullptr, [&] {
        for (int i = 0; i < 4; i++) g->U[i] = upload_mat(G.lat, links[i], eo);
      })) {
    QOP_D3_destroy_G(g);
    return nullptr;
  }
  return g;
}
QOP_F3_GaugeField *QOP_F3_create_G_from_raw(QOP_F_Real *links[], QOP_evenodd_t eo) {
  CHECK_INIT;
  auto *g = new QOP_F3_GaugeField;
  g->lat = G.lat;
  if (!guard_oom(__func__, nullptr, [&] {
        for (int i = 0; i < 4; i++) g->U[i] = upload_mat(G.lat, links[i], eo);
      })) {
    QOP_F3_destroy_G(g);
    return nullptr;
  }
  return g;
}
void QOP_D3_extract_G_to_raw(QOP_D_Real *dest[], QOP_D3_GaugeField *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  for (int i = 0; i < 4; i++) download_mat(dest[i], src->U[i], src->lat, eo);
}
void QOP_F3_extract_G_to_raw(QOP_F_Real *dest[], QOP_F3_GaugeField *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  for (int i = 0; i < 4; i++) download_mat(dest[i], src->U[i], src->lat, eo);
}
void QOP_D3_destroy_G(QOP_D3_GaugeField *g) {
  if (!g) return;
  for (int i = 0; i < 4; i++) if (g->U[i]) QOP_CHECK_HIP(hipFree(g->U[i]));
  delete g;
}
void QOP_F3_destroy_G(QOP_F3_GaugeField *g) {
  if (!g) return;
  for (int i = 0; i < 4; i++) if (g->U[i]) QOP_CHECK_HIP(hipFree(g->U[i]));
  delete g;
}
// common_p.c:199-208, 244-257: the field keeps the caller's arrays; to_raw
// writes it back into them (or into new ones) and destroys it.
extern "C++" {
namespace {
template <class GT, typename T>
T **gauge_to_raw(GT *g, double *const dev[4], int eo, void (*destroy)(GT *)) {
  T **r = g->raw;
  const bool own = !r;
  if (own) {
    r = (T **)malloc(4 * sizeof(T *));
    for (int i = 0; i < 4; i++) r[i] = (T *)calloc((size_t)g->lat->V * 18, sizeof(T));
  }
  sync();
  for (int i = 0; i < 4; i++) download_mat(r[i], dev[i], g->lat, eo);
  destroy(g);
  return r;
}
}  // namespace
}  // extern "C++"
QOP_D3_GaugeField *QOP_D3_convert_G_from_raw(QOP_D_Real *links[], QOP_evenodd_t eo) {
  QOP_D3_GaugeField *g = QOP_D3_create_G_from_raw(links, eo);
  if (g) g->raw = links;
  return g;
}
QOP_F3_GaugeField *QOP_F3_convert_G_from_raw(QOP_F_Real *links[], QOP_evenodd_t eo) {
  QOP_F3_GaugeField *g = QOP_F3_create_G_from_raw(links, eo);
  if (g) g->raw = links;
  return g;
}
QOP_D_Real **QOP_D3_convert_G_to_raw(QOP_D3_GaugeField *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  return gauge_to_raw<QOP_D3_GaugeField, double>(src, src->U, eo, QOP_D3_destroy_G);
}
QOP_F_Real **QOP_F3_convert_G_to_raw(QOP_F3_GaugeField *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  return gauge_to_raw<QOP_F3_GaugeField, float>(src, src->U, eo, QOP_F3_destroy_G);
}
// common_p.c:455-526: sign uses absolute coordinates, bc on the last slice
void QOP_D3_rephase_G(QOP_D3_GaugeField *g, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign) {
  CHECK_INIT;
  rephase_dev(g->lat, g->U, r0, bc, sign, 1, 1, nullptr, 0);
}
void QOP_F3_rephase_G(QOP_F3_GaugeField *g, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign) {
  CHECK_INIT;
  rephase_dev(g->lat, g->U, r0, bc, sign, 1, 1, nullptr, 0);
}

// ---- force fields (common_p.c:80-96, 123-130, 160-169) ------------------------
extern "C++" {
namespace {
template <class FT, typename T>
FT *create_force(T *raw[], int eo, const char *fn) {
  auto *f = new FT;
  f->lat = G.lat;
  if (!guard_oom(fn, nullptr, [&] {
        for (int i = 0; i < 4; i++) f->F[i] = upload_mat(G.lat, raw[i], eo);
      })) {
    for (int i = 0; i < 4; i++) if (f->F[i]) QOP_CHECK_HIP(hipFree(f->F[i]));
    delete f;
    return nullptr;
  }
  return f;
}
template <class FT>
void destroy_force(FT *f) {
  if (!f) return;
  sync();
  for (int i = 0; i < 4; i++) if (f->F[i]) QOP_CHECK_HIP(hipFree(f->F[i]));
  delete f;
}

struct ForceOpts {
  int fnmat_src_min = 2;  // lib/asqtad_force.c:4
  int veclength = 4;
} g_force_opts;

// QOP_asqtad_force_multi (lib/asqtad_force_p.c:41-53 -> asqtad_force_fnmat_p.c:514-546)
template <class GT, class FT, class VT, typename R>
void force_multi(QOP_info_t *info, GT *gauge, FT *force, QOP_asqtad_coeffs_t *c, const R eps[],
                 VT *in_pt[], int nsrc) {
  if (nsrc <= 0) {
    info->final_sec = 0;
    info->final_flop = 0;
    info->status = QOP_SUCCESS;
    return;
  }
  const double coef[6] = {c->one_link, c->three_staple, c->five_staple, c->seven_staple, c->lepage, c->naik};
  std::vector<double> e(eps, eps + nsrc);
  std::vector<const DevVec *> xs(nsrc);
  for (int i = 0; i < nsrc; i++) xs[i] = &in_pt[i]->v;
  double flop = 0;
  const double t0 = wall_time();
  try {
    if (!guard_oom("QOP_asqtad_force_multi", info, [&] {
          asqtad_force_dev(gauge->lat, gauge->U, force->F, coef, e.data(), xs.data(), nsrc, &flop, stream());
        }))
      return;
  } catch (const std::runtime_error &err) {
    fprintf(stderr, "QOP-MI355X: %s\n", err.what());
    info->status = QOP_FAIL;
    return;
  }
  info->final_sec = wall_time() - t0;
  info->final_flop = flop;
  info->status = QOP_SUCCESS;
}
}  // namespace
}  // extern "C++"

QOP_D3_Force *QOP_D3_create_F_from_raw(QOP_D_Real *force[], QOP_evenodd_t eo) {
  CHECK_INIT;
  return create_force<QOP_D3_Force>(force, eo, __func__);
}
QOP_F3_Force *QOP_F3_create_F_from_raw(QOP_F_Real *force[], QOP_evenodd_t eo) {
  CHECK_INIT;
  return create_force<QOP_F3_Force>(force, eo, __func__);
}
void QOP_D3_extract_F_to_raw(QOP_D_Real *dest[], QOP_D3_Force *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  sync();
  for (int i = 0; i < 4; i++) download_mat(dest[i], src->F[i], src->lat, eo);
}
void QOP_F3_extract_F_to_raw(QOP_F_Real *dest[], QOP_F3_Force *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  sync();
  for (int i = 0; i < 4; i++) download_mat(dest[i], src->F[i], src->lat, eo);
}
void QOP_D3_destroy_F(QOP_D3_Force *f) { destroy_force(f); }
void QOP_F3_destroy_F(QOP_F3_Force *f) { destroy_force(f); }
// common_p.c:212-221, 260-273
QOP_D3_Force *QOP_D3_convert_F_from_raw(QOP_D_Real *force[], QOP_evenodd_t eo) {
  QOP_D3_Force *f = QOP_D3_create_F_from_raw(force, eo);
  if (f) f->raw = force;
  return f;
}
QOP_F3_Force *QOP_F3_convert_F_from_raw(QOP_F_Real *force[], QOP_evenodd_t eo) {
  QOP_F3_Force *f = QOP_F3_create_F_from_raw(force, eo);
  if (f) f->raw = force;
  return f;
}
QOP_D_Real **QOP_D3_convert_F_to_raw(QOP_D3_Force *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  return gauge_to_raw<QOP_D3_Force, double>(src, src->F, eo, QOP_D3_destroy_F);
}
QOP_F_Real **QOP_F3_convert_F_to_raw(QOP_F3_Force *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  return gauge_to_raw<QOP_F3_Force, float>(src, src->F, eo, QOP_F3_destroy_F);
}

// lib/asqtad_force.c:22-43: the tags are validated and kept; every nsrc runs
// the fnmat form here.
QOP_status_t QOP_asqtad_force_set_opts(QOP_opt_t opts[], int nopts) {
  CHECK_INIT;
  int fsm = g_force_opts.fnmat_src_min, vl = g_force_opts.veclength;
  for (int i = 0; i < nopts; i++) {
    if (!strcmp(opts[i].tag, "fnmat_src_min")) fsm = (int)opts[i].value;
    if (!strcmp(opts[i].tag, "veclength")) vl = (int)opts[i].value;
  }
  if (fsm < 0 || vl <= 0) return QOP_FAIL;
  g_force_opts.fnmat_src_min = fsm;
  g_force_opts.veclength = vl;
  return QOP_SUCCESS;
}
void QOP_D3_asqtad_force_multi(QOP_info_t *info, QOP_D3_GaugeField *gauge, QOP_D3_Force *force,
                               QOP_asqtad_coeffs_t *coef, QOP_D_Real eps[], QOP_D3_ColorVector *in_pt[],
                               int nsrc) {
  CHECK_INIT;
  force_multi(info, gauge, force, coef, eps, in_pt, nsrc);
}
void QOP_F3_asqtad_force_multi(QOP_info_t *info, QOP_F3_GaugeField *gauge, QOP_F3_Force *force,
                               QOP_asqtad_coeffs_t *coef, QOP_F_Real eps[], QOP_F3_ColorVector *in_pt[],
                               int nsrc) {
  CHECK_INIT;
  force_multi(info, gauge, force, coef, eps, in_pt, nsrc);
}
void QOP_D3_asqtad_force(QOP_info_t *info, QOP_D3_GaugeField *gauge, QOP_D3_Force *force,
                         QOP_asqtad_coeffs_t *coeffs, QOP_D_Real eps, QOP_D3_ColorVector *in_pt) {
  QOP_D3_asqtad_force_multi(info, gauge, force, coeffs, &eps, &in_pt, 1);
}
void QOP_F3_asqtad_force(QOP_info_t *info, QOP_F3_GaugeField *gauge, QOP_F3_Force *force,
                         QOP_asqtad_coeffs_t *coeffs, QOP_F_Real eps, QOP_F3_ColorVector *in_pt) {
  QOP_F3_asqtad_force_multi(info, gauge, force, coeffs, &eps, &in_pt, 1);
}

// ---- links --------------------------------------------------------------------
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_create_L_from_raw(QOP_D_Real *fat[], QOP_D_Real *lng[], QOP_evenodd_t eo) {
  CHECK_INIT;
  auto *f = new QOP_D3_FermionLinksAsqtad;
  f->L.lat = G.lat;
  if (!guard_oom(__func__, nullptr, [&] { build_links_from_raw(f->L, fat, lng, eo); })) {
    QOP_D3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_create_L_from_raw(QOP_F_Real *fat[], QOP_F_Real *lng[], QOP_evenodd_t eo) {
  CHECK_INIT;
  auto *f = new QOP_F3_FermionLinksAsqtad;
  f->L.lat = G.lat;
  if (!guard_oom(__func__, nullptr, [&] { build_links_from_raw(f->L, fat, lng, eo); })) {
    QOP_F3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_create_L_from_G(QOP_info_t *info, QOP_asqtad_coeffs_t *c, QOP_D3_GaugeField *g) {
  CHECK_INIT;
  auto *f = new QOP_D3_FermionLinksAsqtad;
  if (!guard_oom(__func__, info, [&] { fatten(f->L, g->lat, g->U, g->U, c, info); })) {
    QOP_D3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_create_L_from_G(QOP_info_t *info, QOP_asqtad_coeffs_t *c, QOP_F3_GaugeField *g) {
  CHECK_INIT;
  auto *f = new QOP_F3_FermionLinksAsqtad;
  if (!guard_oom(__func__, info, [&] { fatten(f->L, g->lat, g->U, g->U, c, info); })) {
    QOP_F3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_create_L_from_G2(QOP_info_t *info, QOP_asqtad_coeffs_t *c, QOP_D3_GaugeField *gf, QOP_D3_GaugeField *gl) {
  CHECK_INIT;
  auto *f = new QOP_D3_FermionLinksAsqtad;
  if (!guard_oom(__func__, info, [&] { fatten(f->L, gf->lat, gf->U, gl->U, c, info); })) {
    QOP_D3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_create_L_from_G2(QOP_info_t *info, QOP_asqtad_coeffs_t *c, QOP_F3_GaugeField *gf, QOP_F3_GaugeField *gl) {
  CHECK_INIT;
  auto *f = new QOP_F3_FermionLinksAsqtad;
  if (!guard_oom(__func__, info, [&] { fatten(f->L, gf->lat, gf->U, gl->U, c, info); })) {
    QOP_F3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
static void extract_links(Links &L, double *fat[], double *lng[], int eo, float *ffat[], float *flng[]) {
  Lattice *lat = L.lat;
  DevBuf<double> fb[4], lb[4];
  double *fd[4], *ld[4];
  size_t n = (size_t)lat->V * 18 * sizeof(double);
  bool want_long = (lng || flng) && L.nl == 16;
  for (int i = 0; i < 4; i++) {
    dev_alloc(&fb[i].p, n);
    if (want_long) dev_alloc(&lb[i].p, n);
    fd[i] = fb[i].p;
    ld[i] = lb[i].p;
  }
  extract_links64(L, fd, want_long ? ld : nullptr, stream());
  for (int i = 0; i < 4; i++) {
    if (fat) download_mat(fat[i], fd[i], lat, eo);
    if (ffat) download_mat(ffat[i], fd[i], lat, eo);
    if (want_long && lng) download_mat(lng[i], ld[i], lat, eo);
    if (want_long && flng) download_mat(flng[i], ld[i], lat, eo);
  }
}
void QOP_D3_asqtad_extract_L_to_raw(QOP_D_Real *fat[], QOP_D_Real *lng[], QOP_D3_FermionLinksAsqtad *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  guard_oom(__func__, nullptr, [&] { extract_links(src->L, fat, lng, eo, nullptr, nullptr); });
}
void QOP_F3_asqtad_extract_L_to_raw(QOP_F_Real *fat[], QOP_F_Real *lng[], QOP_F3_FermionLinksAsqtad *src, QOP_evenodd_t eo) {
  CHECK_INIT;
  guard_oom(__func__, nullptr, [&] { extract_links(src->L, nullptr, nullptr, eo, fat, lng); });
}
void QOP_D3_asqtad_load_L_from_raw(QOP_D3_FermionLinksAsqtad *f, QOP_D_Real *fat[], QOP_D_Real *lng[], QOP_evenodd_t eo) {
  CHECK_INIT;
  guard_oom(__func__, nullptr, [&] { build_links_from_raw(f->L, fat, lng, eo); });
}
void QOP_D3_asqtad_load_L_from_G(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *f, QOP_asqtad_coeffs_t *c, QOP_D3_GaugeField *g) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] { fatten(f->L, g->lat, g->U, g->U, c, info); });
}
void QOP_F3_asqtad_load_L_from_raw(QOP_F3_FermionLinksAsqtad *f, QOP_F_Real *fat[], QOP_F_Real *lng[], QOP_evenodd_t eo) {
  CHECK_INIT;
  guard_oom(__func__, nullptr, [&] { build_links_from_raw(f->L, fat, lng, eo); });
}
void QOP_F3_asqtad_load_L_from_G(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *f, QOP_asqtad_coeffs_t *c, QOP_F3_GaugeField *g) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] { fatten(f->L, g->lat, g->U, g->U, c, info); });
}
// asqtad_dslash_p.c:878-896: fat links from gFat, Naik links from gLong
void QOP_D3_asqtad_load_L_from_G2(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *f, QOP_asqtad_coeffs_t *c,
                                  QOP_D3_GaugeField *gf, QOP_D3_GaugeField *gl) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] { fatten(f->L, gf->lat, gf->U, gl->U, c, info); });
}
void QOP_F3_asqtad_load_L_from_G2(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *f, QOP_asqtad_coeffs_t *c,
                                  QOP_F3_GaugeField *gf, QOP_F3_GaugeField *gl) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] { fatten(f->L, gf->lat, gf->U, gl->U, c, info); });
}
static void destroy_links(Links &L) {
  sync();
  L.free_all();
}
void QOP_D3_asqtad_destroy_L(QOP_D3_FermionLinksAsqtad *f) {
  if (!f) return;
  destroy_links(f->L);
  delete f;
}
void QOP_F3_asqtad_destroy_L(QOP_F3_FermionLinksAsqtad *f) {
  if (!f) return;
  destroy_links(f->L);
  delete f;
}
// ---- link objects that take the caller's raw arrays -----------------------
// asqtad_dslash_p.c:281-304 -> 152-223: the object takes the arrays and scales
// them in place by 1/2 (:201-206); the device keeps its own double-stored
// copy, the arrays are written back by convert_L_to_raw.
extern "C++" {
namespace {
template <typename T>
void scale_raw(T *const a[4], double s) {
  const size_t n = (size_t)G.lat->V * 18;
  for (int i = 0; i < 4; i++)
    for (size_t e = 0; e < n; e++) a[i][e] = (T)(s * a[i][e]);
}
template <class LT, typename T>
LT *convert_links_from_raw(T *fat[], T *lng[], const char *fn) {
  auto *f = new LT;
  f->L.lat = G.lat;
  if (!guard_oom(fn, nullptr, [&] { build_links_from_raw(f->L, fat, lng, QOP_EVENODD); })) {
    destroy_links(f->L);
    delete f;
    return nullptr;
  }
  scale_raw(fat, 0.5);
  if (lng) scale_raw(lng, 0.5);
  f->raw_fat = fat;
  f->raw_lng = lng;
  return f;
}
template <typename T>
T **alloc_raw4() {
  T **a = (T **)malloc(4 * sizeof(T *));
  for (int i = 0; i < 4; i++) a[i] = (T *)calloc((size_t)G.lat->V * 18, sizeof(T));
  return a;
}
// The current links (user normalisation) go back into the arrays the object
// was converted from, or into new ones; the object is destroyed.  On a device
// allocation failure the object stays valid and *fat = *lng = NULL.
template <class LT, typename T>
void convert_links_to_raw(T ***fat, T ***lng, LT *src, const char *fn) {
  T **rf = src->raw_fat, **rl = src->raw_lng;
  const bool naik = src->L.nl == 16;
  const bool own_f = !rf, own_l = naik && !rl;
  if (own_f) rf = alloc_raw4<T>();
  if (own_l) rl = alloc_raw4<T>();
  bool ok;
  if constexpr (sizeof(T) == 8)
    ok = guard_oom(fn, nullptr, [&] { extract_links(src->L, rf, naik ? rl : nullptr, QOP_EVENODD, nullptr, nullptr); });
  else
    ok = guard_oom(fn, nullptr, [&] { extract_links(src->L, nullptr, nullptr, QOP_EVENODD, rf, naik ? rl : nullptr); });
  if (!ok) {
    auto drop = [](T **a) { for (int i = 0; i < 4; i++) free(a[i]); free(a); };
    if (own_f) drop(rf);
    if (own_l) drop(rl);
    *fat = nullptr;
    if (lng) *lng = nullptr;
    return;
  }
  // long-link arrays of an object reloaded without Naik links: undo the 1/2
  if (!naik && rl) scale_raw(rl, 2.0);
  *fat = rf;
  if (lng) *lng = rl;
  destroy_links(src->L);
  delete src;
}
}  // namespace
}  // extern "C++"

QOP_D3_FermionLinksAsqtad *QOP_D3_asqtad_convert_L_from_raw(QOP_D_Real *fat[], QOP_D_Real *lng[], QOP_evenodd_t) {
  CHECK_INIT;
  return convert_links_from_raw<QOP_D3_FermionLinksAsqtad>(fat, lng, __func__);
}
QOP_F3_FermionLinksAsqtad *QOP_F3_asqtad_convert_L_from_raw(QOP_F_Real *fat[], QOP_F_Real *lng[], QOP_evenodd_t) {
  CHECK_INIT;
  return convert_links_from_raw<QOP_F3_FermionLinksAsqtad>(fat, lng, __func__);
}
void QOP_D3_asqtad_convert_L_to_raw(QOP_D_Real ***fat, QOP_D_Real ***lng, QOP_D3_FermionLinksAsqtad *src, QOP_evenodd_t) {
  CHECK_INIT;
  convert_links_to_raw(fat, lng, src, __func__);
}
void QOP_F3_asqtad_convert_L_to_raw(QOP_F_Real ***fat, QOP_F_Real ***lng, QOP_F3_FermionLinksAsqtad *src, QOP_evenodd_t) {
  CHECK_INIT;
  convert_links_to_raw(fat, lng, src, __func__);
}

// asqtad_dslash_p.c:1385-1482: sign relative to r0, bc on the last slice for
// fat links and the last three slices for long links; then re-double-store.
static void rephase_links(Links &L, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign, const char *fn) {
  Lattice *lat = L.lat;
  guard_oom(fn, nullptr, [&] {
    DevBuf<double> fb[4], lb[4];
    double *fd[4], *ld[4];
    size_t n = (size_t)lat->V * 18 * sizeof(double);
    for (int i = 0; i < 4; i++) {
      dev_alloc(&fb[i].p, n);
      if (L.nl == 16) dev_alloc(&lb[i].p, n);
      fd[i] = fb[i].p;
      ld[i] = lb[i].p;
    }
    extract_links64(L, fd, L.nl == 16 ? ld : nullptr, stream());
    rephase_dev(lat, fd, r0, bc, sign, 1, 3, L.nl == 16 ? ld : nullptr, 1);
    const double *fc[4] = {fd[0], fd[1], fd[2], fd[3]}, *lc[4] = {ld[0], ld[1], ld[2], ld[3]};
    build_links64(L, fc, L.nl == 16 ? lc : nullptr, stream());
  });
}
void QOP_D3_asqtad_rephase_L(QOP_D3_FermionLinksAsqtad *f, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign) {
  CHECK_INIT;
  rephase_links(f->L, r0, bc, sign, __func__);
}
void QOP_F3_asqtad_rephase_L(QOP_F3_FermionLinksAsqtad *f, int *r0, QOP_bc_t *bc, QOP_staggered_sign_t *sign) {
  CHECK_INIT;
  rephase_links(f->L, r0, bc, sign, __func__);
}
// asqtad_dslash_p.c:742-790: the single-precision copy of double links
QOP_F3_FermionLinksAsqtad *QOP_FD3_asqtad_create_L_from_L(QOP_D3_FermionLinksAsqtad *src) {
  CHECK_INIT;
  auto *f = new QOP_F3_FermionLinksAsqtad;
  f->L.lat = src->L.lat;
  f->L.nl = src->L.nl;
  // keep an FP64 master with the FP32-rounded values so both views agree
  const bool ok = guard_oom(__func__, nullptr, [&] {
    LinkSet tmp;
    convert_links(src->L.lat, src->L.l64, tmp, P32, stream());
    f->L.l32 = tmp;  // owned by f from here on
    size_t n = (size_t)2 * src->L.nl * 9 * src->L.lat->Vh;
    dev_alloc(&f->L.l64.data, n * sizeof(double2));
    f->L.l64.prec = P64;
    f->L.l64.nl = src->L.nl;
    // widen FP32 -> FP64 through the vector pack kernels' lincomb path:
    // treat the link array as a batch of n/(3 Vh) parity halves.
    LincombArgs a;
    long Vh = src->L.lat->Vh;
    a.Vh = (int)Vh;
    a.n = (int)(n / (3 * Vh));
    a.out.data = f->L.l64.data;
    a.ostride = 3 * Vh;
    a.v[0].data = f->L.l32.data;
    a.stride[0] = 3 * Vh;
    a.cr[0] = 1.0;
    int pin[1] = {P32};
    launch_lincomb(a, P64, pin, 1, false, stream());
    sync();
  });
  if (!ok) {
    QOP_F3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}
QOP_D3_FermionLinksAsqtad *QOP_DF3_asqtad_create_L_from_L(QOP_F3_FermionLinksAsqtad *src) {
  CHECK_INIT;
  auto *f = new QOP_D3_FermionLinksAsqtad;
  f->L.lat = src->L.lat;
  f->L.nl = src->L.nl;
  const bool ok = guard_oom(__func__, nullptr, [&] {
    size_t n = (size_t)2 * src->L.nl * 9 * src->L.lat->Vh * sizeof(double2);
    dev_alloc(&f->L.l64.data, n);
    f->L.l64.prec = P64;
    f->L.l64.nl = src->L.nl;
    QOP_CHECK_HIP(hipMemcpyAsync(f->L.l64.data, src->L.l64.data, n, hipMemcpyDeviceToDevice, stream()));
    sync();
  });
  if (!ok) {
    QOP_D3_asqtad_destroy_L(f);
    return nullptr;
  }
  return f;
}

// ---- hot path ------------------------------------------------------------------
void QOP_D3_asqtad_dslash(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *f, QOP_D_Real mass,
                          QOP_D3_ColorVector *out, QOP_D3_ColorVector *in, QOP_evenodd_t eo_out,
                          QOP_evenodd_t eo_in) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] {
    dslash(f->L, mass, out->v, in->v, eo_out, eo_in);
    sync();
  });
}
void QOP_F3_asqtad_dslash(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *f, QOP_F_Real mass,
                          QOP_F3_ColorVector *out, QOP_F3_ColorVector *in, QOP_evenodd_t eo_out,
                          QOP_evenodd_t eo_in) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] {
    dslash(f->L, mass, out->v, in->v, eo_out, eo_in);
    sync();
  });
}
// asqtad_dslash_p.c:1140-1191 (asqtad_hop_dir), :1338-1377 (style 1) and
// :1280-1336 (style 0, whose backward term enters with + instead of -)
static void dslash_dir_dev(Links &L, int dir, int fb, double wtfat, double wtlong, DevVec &out,
                           const DevVec &in, int eo_out) {
  QOP_REQUIRE(dir >= 0 && dir < 4, "dslash_dir: dir must be 0..3");
  QOP_REQUIRE(out.prec == in.prec, "dslash_dir: precision mismatch");
  const int prec = out.prec;
  LinkSet &ls = L.get(prec);
  const double bsign = (options().style & 1) ? 1.0 : -1.0;
  std::unique_ptr<TmpVec> tcopy;  // the reference gathers from a copy of the source too
  const DevVec *src = &in;
  if (out.data == in.data) {
    tcopy.reset(new TmpVec(in.lat, in.prec));
    copy(tcopy->v, in, 2);
    src = &tcopy->v;
  }
  for (int p = 0; p < 2; p++) {
    if (eo_out != 2 && eo_out != p) continue;
    launch_hop_dir(L.lat, ls, prec, p, dir, fb, wtfat, wtlong, bsign, out.half(p), src->half(1 - p), stream());
  }
  sync();
}
void QOP_D3_asqtad_dslash_dir(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *f, int dir, int fb, double wtfat,
                              double wtlong, QOP_D3_ColorVector *out, QOP_D3_ColorVector *in,
                              QOP_evenodd_t eo_out) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] { dslash_dir_dev(f->L, dir, fb, wtfat, wtlong, out->v, in->v, eo_out); });
}
void QOP_F3_asqtad_dslash_dir(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *f, int dir, int fb, double wtfat,
                              double wtlong, QOP_F3_ColorVector *out, QOP_F3_ColorVector *in,
                              QOP_evenodd_t eo_out) {
  CHECK_INIT;
  guard_oom(__func__, info, [&] { dslash_dir_dev(f->L, dir, fb, wtfat, wtlong, out->v, in->v, eo_out); });
}
void QOP_D3_asqtad_diaginv(QOP_info_t *, QOP_D3_FermionLinksAsqtad *, QOP_D_Real mass,
                           QOP_D3_ColorVector *out, QOP_D3_ColorVector *in, QOP_evenodd_t eo) {
  CHECK_INIT;
  scal(out->v, Cplx{1.0 / mass, 0}, in->v, eo);
  sync();
}
void QOP_F3_asqtad_diaginv(QOP_info_t *, QOP_F3_FermionLinksAsqtad *, QOP_F_Real mass,
                           QOP_F3_ColorVector *out, QOP_F3_ColorVector *in, QOP_evenodd_t eo) {
  CHECK_INIT;
  scal(out->v, Cplx{(double)(1.0f / mass), 0}, in->v, eo);
  sync();
}

}  // extern "C"

// solver dispatch: reference algorithm, or this library's mixed precision
namespace qop {
void invert_dispatch(Links &L, InvArg *ia, ResArg *ra, double mass, DevVec &out, const DevVec &in,
                     double *flop, double *sec);
void invert_multi_dispatch(Links &L, InvArg *ia, ResArg **ra, const double *masses, const int *nmass,
                           DevVec **out, DevVec *const *in, int nsrc, double *flop, double *sec);
}  // namespace qop

static InvArg to_ia(const QOP_invert_arg_t *a) {
  return InvArg{a->mixed_rsq, a->max_iter, a->restart, a->max_restarts, (int)a->evenodd};
}
static ResArg to_ra(const QOP_resid_arg_t *r) {
  return ResArg{r->rsqmin, r->final_rsq, r->relmin, r->final_rel, r->final_iter, r->final_restart};
}
static void from_ra(QOP_resid_arg_t *r, const ResArg &a) {
  r->final_rsq = a.final_rsq;
  r->final_rel = a.final_rel;
  r->final_iter = a.final_iter;
  r->final_restart = a.final_restart;
}

extern "C" {

void QOP_D3_asqtad_invert(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *f, QOP_invert_arg_t *inv_arg,
                          QOP_resid_arg_t *res_arg, QOP_D_Real mass, QOP_D3_ColorVector *out,
                          QOP_D3_ColorVector *in) {
  CHECK_INIT;
  InvArg ia = to_ia(inv_arg);
  ResArg ra = to_ra(res_arg);
  double flop = 0, sec = 0;
  if (!guard_oom(__func__, info, [&] { invert_dispatch(f->L, &ia, &ra, mass, out->v, in->v, &flop, &sec); }))
    return;
  from_ra(res_arg, ra);
  if (info) { info->final_sec = sec; info->final_flop = flop; info->status = QOP_SUCCESS; }
}

// Single precision: the solve iterates on the FP32 vectors with the FP32
// links, as lib/pcstubs/asqtad_invert_f3.c instantiates lib/asqtad_invert_p.c
// with QOP_Precision 'F' (scalars and reductions stay FP64).
void QOP_F3_asqtad_invert(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *f, QOP_invert_arg_t *inv_arg,
                          QOP_resid_arg_t *res_arg, QOP_F_Real mass, QOP_F3_ColorVector *out,
                          QOP_F3_ColorVector *in) {
  CHECK_INIT;
  InvArg ia = to_ia(inv_arg);
  ResArg ra = to_ra(res_arg);
  double flop = 0, sec = 0;
  if (!guard_oom(__func__, info, [&] { invert_dispatch(f->L, &ia, &ra, mass, out->v, in->v, &flop, &sec); }))
    return;
  from_ra(res_arg, ra);
  if (info) { info->final_sec = sec; info->final_flop = flop; info->status = QOP_SUCCESS; }
}

void QOP_D3_asqtad_invert_multi(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *f, QOP_invert_arg_t *inv_arg,
                                QOP_resid_arg_t **res_arg[], QOP_D_Real *masses[], int nmass[],
                                QOP_D3_ColorVector **out_pt[], QOP_D3_ColorVector *in_pt[], int nsrc) {
  CHECK_INIT;
  InvArg ia = to_ia(inv_arg);
  int nsolve = 0;
  for (int i = 0; i < nsrc; i++) nsolve += nmass[i];
  std::vector<ResArg> ra(nsolve);
  std::vector<ResArg *> rap(nsolve);
  std::vector<double> ms(nsolve);
  std::vector<DevVec *> out(nsolve), in(nsrc);
  int k = 0;
  for (int i = 0; i < nsrc; i++) {
    in[i] = &in_pt[i]->v;
    for (int j = 0; j < nmass[i]; j++, k++) {
      ra[k] = to_ra(res_arg[i][j]);
      rap[k] = &ra[k];
      ms[k] = masses[i][j];
      out[k] = &out_pt[i][j]->v;
    }
  }
  double flop = 0, sec = 0;
  if (!guard_oom(__func__, info, [&] {
        invert_multi_dispatch(f->L, &ia, rap.data(), ms.data(), nmass, out.data(), in.data(), nsrc, &flop,
                              &sec);
      }))
    return;
  k = 0;
  for (int i = 0; i < nsrc; i++)
    for (int j = 0; j < nmass[i]; j++, k++) from_ra(res_arg[i][j], ra[k]);
  if (info) { info->final_sec = sec; info->final_flop = flop; info->status = QOP_SUCCESS; }
}

void QOP_F3_asqtad_invert_multi(QOP_info_t *info, QOP_F3_FermionLinksAsqtad *f, QOP_invert_arg_t *inv_arg,
                                QOP_resid_arg_t **res_arg[], QOP_F_Real *masses[], int nmass[],
                                QOP_F3_ColorVector **out_pt[], QOP_F3_ColorVector *in_pt[], int nsrc) {
  CHECK_INIT;
  InvArg ia = to_ia(inv_arg);
  int nsolve = 0;
  for (int i = 0; i < nsrc; i++) nsolve += nmass[i];
  std::vector<ResArg> ra(nsolve);
  std::vector<ResArg *> rap(nsolve);
  std::vector<double> ms(nsolve);
  std::vector<DevVec *> out(nsolve), in(nsrc);
  int k = 0;
  for (int i = 0; i < nsrc; i++) {
    in[i] = &in_pt[i]->v;
    for (int j = 0; j < nmass[i]; j++, k++) {
      ra[k] = to_ra(res_arg[i][j]);
      rap[k] = &ra[k];
      ms[k] = masses[i][j];
      out[k] = &out_pt[i][j]->v;
    }
  }
  double flop = 0, sec = 0;
  if (!guard_oom(__func__, info, [&] {
        invert_multi_dispatch(f->L, &ia, rap.data(), ms.data(), nmass, out.data(), in.data(), nsrc, &flop,
                              &sec);
      }))
    return;
  k = 0;
  for (int i = 0; i < nsrc; i++)
    for (int j = 0; j < nmass[i]; j++, k++) from_ra(res_arg[i][j], ra[k]);
  if (info) { info->final_sec = sec; info->final_flop = flop; info->status = QOP_SUCCESS; }
}

void *QOPX_host_alloc(unsigned long long bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
void QOPX_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

int QOPX_device_ok(void) {
  int nd = 0;
  return hipGetDeviceCount(&nd) == hipSuccess && nd > 0;
}
const char *QOPX_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
