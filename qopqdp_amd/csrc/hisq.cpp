// QOP_{D3,F3}_hisq_create_L_from_G and friends: HISQ fermion links as the
// reference builds them (lib/hisq_links_p.c:62-256), all on the device:
//   fat7 smearing of the (rephased) gauge field, V = 2 x (internal 1/2 link),
//   U(3) projection W = V (V^dagger V)^-1/2 of every link (hisq.hip),
//   asqtad smearing of W with the level-2 coefficients (fat + Naik), and for
//   n_naiks > 1 the epsilon-Naik sets fn[i] = L2 + eps_i L3 with L3 the
//   "difference" coefficients.
// Options as lib/hisq_links.c:4-77 (QOP_hisq_links_set_opts).  With want_aux
// (the default) the auxiliary links the HISQ force needs stay on the device:
// U (the rephased gauge field), V (fat7) and W (its U(3) projection; Y = W
// for U(3) projection), raw FP64.
// QOP_{D3,F3}_hisq_force_multi (lib/hisq_force_p.c:63-83) and
// QOP_hisq_force_set_opts (lib/hisq_force.c:4-52) are here too.
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "handles.h"
#include "hisq.h"
#include "solver.h"

using namespace qop;

struct QOP_D3_FermionLinksHisq_struct {
  int n_naiks = 1;
  std::vector<QOP_D3_FermionLinksAsqtad *> fn;
  QOP_D3_FermionLinksAsqtad *fn_deps = nullptr;
  qop::Lattice *lat = nullptr;
  double *aux[3][4] = {};  // U, V, W (want_aux)
};
struct QOP_F3_FermionLinksHisq_struct {
  int n_naiks = 1;
  std::vector<QOP_F3_FermionLinksAsqtad *> fn;
  QOP_F3_FermionLinksAsqtad *fn_deps = nullptr;
  qop::Lattice *lat = nullptr;
  double *aux[3][4] = {};  // U, V, W (want_aux)
};

namespace {

struct HisqOpts {
  int want_deps = 0, want_aux = 1;
  int reunit_allow_svd = 1, reunit_svd_only = 0;
  double reunit_svd_rel_error = 1e-8, reunit_svd_abs_error = 1e-8;
  int svd_values_info = 1, use_fat7_lepage = 0;
} g_hisq;

double now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// raw FP64 device matrices [4][V][18], freed on scope exit
struct Raw4 {
  double *m[4] = {nullptr, nullptr, nullptr, nullptr};
  Raw4() = default;
  Raw4(const Raw4 &) = delete;
  Raw4 &operator=(const Raw4 &) = delete;
  ~Raw4() { free(); }
  void alloc(const Lattice *lat) {
    for (int i = 0; i < 4; i++) dev_alloc(&m[i], (size_t)lat->V * 18 * sizeof(double));
  }
  void free() {
    for (int i = 0; i < 4; i++)
      if (m[i]) QOP_CHECK_HIP(hipFree(m[i]));
    for (int i = 0; i < 4; i++) m[i] = nullptr;
  }
};

void coef6(const QOP_asqtad_coeffs_t &c, double out[6]) {
  out[0] = c.one_link; out[1] = c.three_staple; out[2] = c.five_staple;
  out[3] = c.seven_staple; out[4] = c.lepage; out[5] = c.naik;
}

bool has_staples(const QOP_asqtad_coeffs_t &c) {
  return c.three_staple != 0 || c.five_staple != 0 || c.seven_staple != 0 || c.lepage != 0;
}

// asqtad smearing flop count of one create_L_from_G (as the link builder reports it)
double smear_flops(const QOP_asqtad_coeffs_t &c, const Lattice *lat) {
  return (has_staples(c) || c.naik != 0) ? 61632.0 * lat->V : 0.0;
}

template <typename LinksT>
LinksT *links_from_raw_dev(Lattice *lat, const Raw4 &fat, const Raw4 *lng) {
  auto *f = new LinksT;
  f->L.lat = lat;
  const double *fp[4] = {fat.m[0], fat.m[1], fat.m[2], fat.m[3]};
  const double *lp[4] = {nullptr, nullptr, nullptr, nullptr};
  if (lng)
    for (int i = 0; i < 4; i++) lp[i] = lng->m[i];
  try {
    build_links64(f->L, fp, lng ? lp : nullptr, stream());
  } catch (const DeviceOutOfMemory &) {
    f->L.free_all();
    delete f;
    throw;
  }
  sync();
  return f;
}

template <typename HisqT, typename LinksT, typename GaugeT>
HisqT *hisq_create(QOP_info_t *info, QOP_hisq_coeffs_t *c, GaugeT *gauge) {
  QOP_REQUIRE(gauge && c, "hisq_create_L_from_G: null argument");
  if (c->ugroup == QOP_UNITARIZE_SU3) {
    // lib/hisq_links_p.c:162-170
    printf("QOP_hisq_create_L_from_G: SU(3) projection not supported for now\n");
    exit(1);
  }
  const int n_naiks = c->n_naiks < 1 ? 1 : c->n_naiks;
  QOP_REQUIRE(n_naiks <= QOP_MAX_NAIK, "hisq_create_L_from_G: n_naiks > QOP_MAX_NAIK");
  Lattice *lat = gauge->lat;
  hipStream_t st = stream();
  const double t0 = now();
  double flops = 0;

  QOP_asqtad_coeffs_t a1 = {c->fat7_one_link, c->fat7_three_staple, c->fat7_five_staple,
                            c->fat7_seven_staple, g_hisq.use_fat7_lepage ? c->fat7_lepage : 0.0, 0.0};
  QOP_asqtad_coeffs_t a2 = {c->asqtad_one_link, c->asqtad_three_staple, c->asqtad_five_staple,
                            c->asqtad_seven_staple, c->asqtad_lepage, c->asqtad_naik};
  QOP_asqtad_coeffs_t a3 = {c->difference_one_link, 0, 0, 0, 0, c->difference_naik};

  // fat7 stage (hisq_links_p.c:118-128); the raw smeared link is the
  // reference's 2 x internal link.  Without want_aux it is projected in place.
  const bool aux = g_hisq.want_aux != 0;
  Raw4 W, Vf, Uc;
  W.alloc(lat);
  if (aux) {
    Vf.alloc(lat);
    Uc.alloc(lat);
    for (int i = 0; i < 4; i++)
      QOP_CHECK_HIP(hipMemcpyAsync(Uc.m[i], gauge->U[i], (size_t)lat->V * 18 * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
  }
  double **V7 = aux ? Vf.m : W.m;
  {
    double k[6];
    coef6(a1, k);
    const double *g[4] = {gauge->U[0], gauge->U[1], gauge->U[2], gauge->U[3]};
    launch_fatten(lat, g, g, k, V7, nullptr, st);
    flops += smear_flops(a1, lat) + 198.0 * lat->V;
  }
  // U(3) projection in place (hisq_links_p.c:139-150)
  int nsvd = 0;
  {
    U3Args u;
    u.V = lat->V;
    for (int i = 0; i < 4; i++) {
      u.in[i] = V7[i];
      u.out[i] = W.m[i];
    }
    u.allow_svd = g_hisq.reunit_allow_svd;
    u.svd_only = g_hisq.reunit_svd_only;
    u.rel_error = g_hisq.reunit_svd_rel_error;
    u.abs_error = g_hisq.reunit_svd_abs_error;
    const long nw = u3_nwaves(lat);
    dev_alloc(&u.counts, 2 * nw * sizeof(int));
    launch_u3_project(u, st);
    std::vector<int> cnt(2 * nw);
    QOP_CHECK_HIP(hipMemcpyAsync(cnt.data(), u.counts, 2 * nw * sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    QOP_CHECK_HIP(hipFree(u.counts));
    long bad = 0;
    for (long i = 0; i < nw; i++) {
      nsvd += cnt[2 * i];
      bad += cnt[2 * i + 1];
    }
    QOP_REQUIRE(bad == 0, "hisq_create_L_from_G: singular fat7 link, no U(3) projection");
    // flops of the analytic path per link (hisq_reunit_utilities_p.c:97-297)
    flops += 795.0 * 4 * lat->V + 202.0 * nsvd;
  }
  auto *h = new HisqT;
  // frees h and the link sets built so far if a device allocation fails
  struct Owner {
    HisqT *h;
    ~Owner() {
      if (!h) return;
      sync();
      for (auto *f : h->fn)
        if (f) { f->L.free_all(); delete f; }
      if (h->fn_deps) { h->fn_deps->L.free_all(); delete h->fn_deps; }
      for (auto &a : h->aux)
        for (auto *p : a)
          if (p) QOP_CHECK_HIP(hipFree(p));
      delete h;
    }
  } owner{h};
  h->lat = lat;
  h->n_naiks = n_naiks;
  h->fn.assign(n_naiks, nullptr);
  const double *w[4] = {W.m[0], W.m[1], W.m[2], W.m[3]};
  Raw4 fat2, lng2;
  fat2.alloc(lat);
  const bool naik2 = a2.naik != 0;
  if (naik2) lng2.alloc(lat);
  {
    double k[6];
    coef6(a2, k);
    launch_fatten(lat, w, w, k, fat2.m, naik2 ? lng2.m : nullptr, st);
    flops += smear_flops(a2, lat);
  }
  if (n_naiks > 1 || g_hisq.want_deps) {
    // the "difference" set (hisq_links_p.c:195-221, 241-247)
    Raw4 fat3, lng3;
    fat3.alloc(lat);
    const bool naik3 = a3.naik != 0;
    if (naik3) lng3.alloc(lat);
    double k[6];
    coef6(a3, k);
    launch_fatten(lat, w, w, k, fat3.m, naik3 ? lng3.m : nullptr, st);
    flops += smear_flops(a3, lat);
    if (g_hisq.want_deps) h->fn_deps = links_from_raw_dev<LinksT>(lat, fat3, naik3 ? &lng3 : nullptr);
    const long n = lat->V * 18;
    Raw4 fi, li;
    fi.alloc(lat);
    if (naik2 || naik3) li.alloc(lat);
    for (int in = 1; in < n_naiks; in++) {
      const double eps = c->eps_naik[in];
      for (int mu = 0; mu < 4; mu++) {
        launch_axpy_raw(fi.m[mu], fat2.m[mu], eps, fat3.m[mu], n, st);
        if (naik2 && naik3) launch_axpy_raw(li.m[mu], lng2.m[mu], eps, lng3.m[mu], n, st);
        else if (naik2) QOP_CHECK_HIP(hipMemcpyAsync(li.m[mu], lng2.m[mu], n * sizeof(double), hipMemcpyDeviceToDevice, st));
        else if (naik3) launch_axpy_raw(li.m[mu], nullptr, eps, lng3.m[mu], n, st);
      }
      h->fn[in] = links_from_raw_dev<LinksT>(lat, fi, (naik2 || naik3) ? &li : nullptr);
      flops += (naik2 || naik3 ? 8.0 : 4.0) * 18 * lat->V * 2;
    }
    fi.free();
    li.free();
    fat3.free();
    lng3.free();
  }
  h->fn[0] = links_from_raw_dev<LinksT>(lat, fat2, naik2 ? &lng2 : nullptr);
  fat2.free();
  lng2.free();
  if (aux) {
    for (int i = 0; i < 4; i++) {
      h->aux[0][i] = Uc.m[i];
      h->aux[1][i] = Vf.m[i];
      h->aux[2][i] = W.m[i];
      Uc.m[i] = Vf.m[i] = W.m[i] = nullptr;
    }
  }
  W.free();
  sync();
  if (info) {
    info->final_flop = flops;
    info->final_sec = now() - t0;
    info->status = QOP_SUCCESS;
    info->count1 += nsvd;  // QOP_info_hisq_svd_counter
  }
  owner.h = nullptr;
  return h;
}

template <typename HisqT, typename DestroyL>
void hisq_destroy(HisqT *h, DestroyL destroy) {
  if (!h) return;
  for (auto *f : h->fn)
    if (f) destroy(f);
  if (h->fn_deps) destroy(h->fn_deps);
  sync();
  for (auto &a : h->aux)
    for (auto *p : a)
      if (p) QOP_CHECK_HIP(hipFree(p));
  delete h;
}

struct HisqForceOpts {
  int fnmat_src_min = 1, veclength = 4;  // lib/hisq_force.c:4-9
  double force_filter = 5e-5;
} g_hforce;

// QOP_hisq_force_multi (lib/hisq_force_p.c:63-83, hisq_force_fnmat_p.c:193-231)
template <class HisqT, class FT, class VT, typename R>
void hisq_force_multi(QOP_info_t *info, HisqT *flh, FT *force, QOP_hisq_coeffs_t *c, const R eps[],
                      VT *in_pt[], int *n_orders_naik) {
  const int n_naiks = c->n_naiks < 1 ? 1 : c->n_naiks;
  int nterms = 0;
  for (int k = 0; k < n_naiks; k++) nterms += n_orders_naik[k];
  if (!flh->aux[0][0]) {
    fprintf(stderr, "QOP-MI355X: hisq_force_multi: the HISQ links were built without want_aux\n");
    info->status = QOP_FAIL;
    return;
  }
  if (c->umethod != QOP_UNITARIZE_RATIONAL) {
    // lib/hisq_force_fnmat_p.c:130-169: NONE differentiates the fat7 links
    // directly; only the rational (U(3) projection) form is built here
    fprintf(stderr, "QOP-MI355X: hisq_force_multi: unitarize method %d not supported\n", (int)c->umethod);
    info->status = QOP_FAIL;
    return;
  }
  if (nterms <= 0) {
    info->final_sec = info->final_flop = 0;
    info->status = QOP_SUCCESS;
    return;
  }
  const double fat7[6] = {c->fat7_one_link, c->fat7_three_staple, c->fat7_five_staple,
                          c->fat7_seven_staple, g_hisq.use_fat7_lepage ? c->fat7_lepage : 0.0, 0.0};
  const double lev2[6] = {c->asqtad_one_link, c->asqtad_three_staple, c->asqtad_five_staple,
                          c->asqtad_seven_staple, c->asqtad_lepage, c->asqtad_naik};
  const double diff[6] = {c->difference_one_link, 0, 0, 0, 0, c->difference_naik};
  std::vector<double> e(eps, eps + nterms);
  std::vector<const DevVec *> xs(nterms);
  for (int i = 0; i < nterms; i++) xs[i] = &in_pt[i]->v;
  double flop = 0;
  int nfilt = 0;
  const double t0 = now();
  try {
    if (!guard_oom("QOP_hisq_force_multi", info, [&] {
          hisq_force_dev(flh->lat, flh->aux[0], flh->aux[1], flh->aux[2], force->F, fat7, lev2, diff,
                         n_naiks, c->eps_naik, e.data(), xs.data(), n_orders_naik,
                         g_hforce.force_filter, &flop, &nfilt, stream());
        }))
      return;
  } catch (const std::runtime_error &err) {
    fprintf(stderr, "QOP-MI355X: %s\n", err.what());
    info->status = QOP_FAIL;
    return;
  }
  info->final_sec = now() - t0;
  info->final_flop = flop;
  info->count2 += nfilt;  // QOP_info_hisq_force_filter_counter
  info->status = QOP_SUCCESS;
}

}  // namespace

extern "C" {

// lib/hisq_links.c:31-77
QOP_status_t QOP_hisq_links_set_opts(QOP_opt_t opts[], int nopts) {
  for (int i = 0; i < nopts; i++) {
    const char *t = opts[i].tag;
    const double v = opts[i].value;
    if (!strcmp(t, "want_deps")) g_hisq.want_deps = (int)v;
    else if (!strcmp(t, "want_aux")) g_hisq.want_aux = (int)v;
    else if (!strcmp(t, "reunit_allow_svd")) g_hisq.reunit_allow_svd = (int)v;
    else if (!strcmp(t, "reunit_svd_only")) g_hisq.reunit_svd_only = (int)v;
    else if (!strcmp(t, "reunit_svd_rel_error")) g_hisq.reunit_svd_rel_error = v;
    else if (!strcmp(t, "reunit_svd_abs_error")) g_hisq.reunit_svd_abs_error = v;
    else if (!strcmp(t, "svd_values_info")) g_hisq.svd_values_info = (int)v;
    else if (!strcmp(t, "use_fat7_lepage")) g_hisq.use_fat7_lepage = (int)v;
  }
  return QOP_SUCCESS;
}

QOP_D3_FermionLinksHisq *QOP_D3_hisq_create_L_from_G(QOP_info_t *info, QOP_hisq_coeffs_t *coeffs,
                                                     QOP_D3_GaugeField *gauge) {
  QOP_D3_FermionLinksHisq *h = nullptr;
  guard_oom(__func__, info, [&] {
    h = hisq_create<QOP_D3_FermionLinksHisq, QOP_D3_FermionLinksAsqtad>(info, coeffs, gauge);
  });
  return h;
}
QOP_F3_FermionLinksHisq *QOP_F3_hisq_create_L_from_G(QOP_info_t *info, QOP_hisq_coeffs_t *coeffs,
                                                     QOP_F3_GaugeField *gauge) {
  QOP_F3_FermionLinksHisq *h = nullptr;
  guard_oom(__func__, info, [&] {
    h = hisq_create<QOP_F3_FermionLinksHisq, QOP_F3_FermionLinksAsqtad>(info, coeffs, gauge);
  });
  return h;
}
void QOP_D3_hisq_destroy_L(QOP_D3_FermionLinksHisq *h) { hisq_destroy(h, QOP_D3_asqtad_destroy_L); }
void QOP_F3_hisq_destroy_L(QOP_F3_FermionLinksHisq *h) { hisq_destroy(h, QOP_F3_asqtad_destroy_L); }
QOP_D3_FermionLinksAsqtad **QOP_D3_get_asqtad_links_from_hisq(QOP_D3_FermionLinksHisq *h) {
  return h ? h->fn.data() : nullptr;
}
QOP_F3_FermionLinksAsqtad **QOP_F3_get_asqtad_links_from_hisq(QOP_F3_FermionLinksHisq *h) {
  return h ? h->fn.data() : nullptr;
}
QOP_D3_FermionLinksAsqtad *QOP_D3_get_asqtad_deps_links_from_hisq(QOP_D3_FermionLinksHisq *h) {
  return h ? h->fn_deps : nullptr;
}
QOP_F3_FermionLinksAsqtad *QOP_F3_get_asqtad_deps_links_from_hisq(QOP_F3_FermionLinksHisq *h) {
  return h ? h->fn_deps : nullptr;
}

// lib/hisq_force.c:29-52: every nsrc takes the fnmat form here
QOP_status_t QOP_hisq_force_set_opts(QOP_opt_t opts[], int nopts) {
  int fsm = g_hforce.fnmat_src_min, vl = g_hforce.veclength;
  double ff = g_hforce.force_filter;
  for (int i = 0; i < nopts; i++) {
    if (!strcmp(opts[i].tag, "fnmat_src_min")) fsm = (int)opts[i].value;
    if (!strcmp(opts[i].tag, "veclength")) vl = (int)opts[i].value;
    if (!strcmp(opts[i].tag, "force_filter")) ff = opts[i].value;
  }
  if (fsm < 0 || vl <= 0 || !(ff >= 0)) return QOP_FAIL;
  g_hforce.fnmat_src_min = fsm;
  g_hforce.veclength = vl;
  g_hforce.force_filter = ff;
  return QOP_SUCCESS;
}
void QOP_D3_hisq_force_multi(QOP_info_t *info, QOP_D3_FermionLinksHisq *flh, QOP_D3_Force *force,
                             QOP_hisq_coeffs_t *coef, QOP_D_Real eps[], QOP_D3_ColorVector *in_pt[],
                             int *n_orders_naik) {
  hisq_force_multi(info, flh, force, coef, eps, in_pt, n_orders_naik);
}
void QOP_F3_hisq_force_multi(QOP_info_t *info, QOP_F3_FermionLinksHisq *flh, QOP_F3_Force *force,
                             QOP_hisq_coeffs_t *coef, QOP_F_Real eps[], QOP_F3_ColorVector *in_pt[],
                             int *n_orders_naik) {
  hisq_force_multi(info, flh, force, coef, eps, in_pt, n_orders_naik);
}

}  // extern "C"
