// Field operations on device vectors (see ops.h).
#include <memory>

#include "ops.h"
#include "solver.h"

namespace qop {

namespace {
// Slots of a subset: even -> half 0, odd -> half 1, both -> halves 0,1.
struct Sub {
  int first, n;
};
Sub subset(int eo) {
  if (eo == 0) return {0, 1};
  if (eo == 1) return {1, 1};
  return {0, 2};
}
}  // namespace

void set_hop_links(HopArgs &a, Links &L, int prec, int out_parity) {
  Lattice *lat = L.lat;
  if ((options().style & 1) == 0 && prec != P16) {
    LinkSet &s = L.get_single(prec);
    a.single = 1;
    a.links = s.parity_data(lat, out_parity);
    a.links_bck = s.parity_data(lat, 1 - out_parity);
    a.lscale = nullptr;
  } else {
    LinkSet &ls = L.get(prec);
    a.single = 0;
    a.links = ls.parity_data(lat, out_parity);
    a.lscale = ls.parity_scale(lat, out_parity);
  }
}

void hop(Links &L, int prec, DevVec &dst, int out_parity, const DevVec &src,
         double alpha, double beta, const DevVec *x, int red, double *dev_result,
         hipStream_t st, const int *status) {
  Lattice *lat = L.lat;
  QOP_REQUIRE(dst.prec == prec && src.prec == prec && (!x || x->prec == prec),
              "hop: precision mismatch");
  HopArgs a;
  a.g = lat->g;
  a.parity = out_parity;
  set_hop_links(a, L, prec, out_parity);
  a.src = src.half(1 - out_parity);
  a.dst = dst.half(out_parity);
  if (x) a.x = x->half(out_parity);
  a.n = 1;
  a.alpha = alpha;
  a.beta = beta;
  a.status = status;
  int nch = hop_partial_blocks(lat);
  if (red != RED_NONE) a.partial = red_partial(nch);
  launch_hop_args(a, prec, L.nl, 1, 256, red, x != nullptr, st);
  if (red != RED_NONE) launch_finalize(a.partial, nch, 1, dev_result, st, status);
}

// lib/asqtad_dslash_p.c:1048-1092: which parities get the hop, the mass term,
// or both.  The mass term is out = mass*in (+ hop).
void dslash(Links &L, double mass, DevVec &out, const DevVec &in, int eo_out, int eo_in) {
  hipStream_t st = stream();
  const int prec = out.prec;
  // The hop reads `in` on the opposite parity while writing `out`; when they
  // alias, use a copy of the source (the reference copies it too, :1003-1007).
  std::unique_ptr<TmpVec> tcopy;
  const DevVec *src = &in;
  if (out.data == in.data) {
    tcopy.reset(new TmpVec(in.lat, in.prec));
    copy(tcopy->v, in, 2);
    src = &tcopy->v;
  }
  auto mass_only = [&](int par) {
    Cplx c{mass, 0};
    scal(out, c, *src, par);
  };
  auto hop_only = [&](int par) { hop(L, prec, out, par, *src, 1.0, 0.0, nullptr, RED_NONE, nullptr, st); };
  auto hop_mass = [&](int par) { hop(L, prec, out, par, *src, 1.0, mass, src, RED_NONE, nullptr, st); };
  if (eo_in == eo_out) {
    if (eo_out == 2) { hop_mass(0); hop_mass(1); }
    else mass_only(eo_out);
  } else {
    if (eo_out == 0 || eo_out == 2) {
      if (eo_in == 1) hop_only(0);
      else if (eo_in == 0) mass_only(0);
      else hop_mass(0);
    }
    if (eo_out == 1 || eo_out == 2) {
      if (eo_in == 0) hop_only(1);
      else if (eo_in == 1) mass_only(1);
      else hop_mass(1);
    }
  }
  if (tcopy) {
    sync();
    tcopy.reset();
  }
}

static double *reduce_sync(int mode, const DevVec &a, const DevVec *b, int eo, int &nval) {
  Lattice *lat = a.lat;
  Sub s = subset(eo);
  DotArgs d;
  d.Vh = (int)lat->Vh;
  d.n = s.n;
  d.u = a.half(s.first);
  d.ustride = 3 * lat->Vh;
  if (b) { d.v = b->half(s.first); d.vstride = 3 * lat->Vh; }
  d.nstride = lat->Vh;
  int nch = hop_partial_blocks(lat);
  nval = (mode == 1 ? 2 : 1) * s.n;
  d.partial = red_partial((size_t)nval * nch);
  launch_dot(d, a.prec, b ? b->prec : a.prec, mode, stream());
  double *res = red_result(nval);
  launch_finalize(d.partial, nch, nval, res, stream(), nullptr);
  double *hr = red_host(nval);
  QOP_CHECK_HIP(hipMemcpyAsync(hr, res, nval * sizeof(double), hipMemcpyDeviceToHost, stream()));
  sync();
  return hr;
}

double norm2(const DevVec &v, int eo) {
  int n;
  double *r = reduce_sync(0, v, nullptr, eo, n);
  double s = 0;
  for (int i = 0; i < n; i++) s += r[i];
  return s;
}

Cplx cdot(const DevVec &a, const DevVec &b, int eo) {
  int n;
  double *r = reduce_sync(1, a, &b, eo, n);
  Cplx c{0, 0};
  for (int i = 0; i < n; i += 2) { c.re += r[i]; c.im += r[i + 1]; }
  return c;
}

double redot(const DevVec &a, const DevVec &b, int eo) { return cdot(a, b, eo).re; }

double relnorm2(const DevVec &r, const DevVec &x, int eo) {
  int n;
  double *p = reduce_sync(2, r, &x, eo, n);
  double s = 0;
  for (int i = 0; i < n; i++) s += p[i];
  long count = eo == 2 ? r.lat->V : r.lat->Vh;
  return s / (double)count;
}

void lincomb(DevVec &out, int eo, int nin, const Cplx *c, const DevVec *const *v,
             hipStream_t st) {
  if (!st) st = stream();
  Lattice *lat = out.lat;
  Sub s = subset(eo);
  LincombArgs a;
  a.Vh = (int)lat->Vh;
  a.n = s.n;
  a.out = out.half(s.first);
  a.ostride = 3 * lat->Vh;
  a.nstride = lat->Vh;
  int pin[3];
  for (int i = 0; i < nin; i++) {
    a.v[i] = v[i]->half(s.first);
    a.stride[i] = 3 * lat->Vh;
    a.cr[i] = c[i].re;
    a.ci[i] = c[i].im;
    pin[i] = v[i]->prec;
  }
  launch_lincomb(a, out.prec, pin, nin, false, st);
}

void copy(DevVec &d, const DevVec &s, int eo) {
  Cplx one{1, 0};
  const DevVec *v[1] = {&s};
  lincomb(d, eo, 1, &one, v);
}

void zero(DevVec &d, int eo) {
  Sub s = subset(eo);
  const long Vh = d.lat->Vh;
  VecView h = d.half(s.first);
  QOP_CHECK_HIP(hipMemsetAsync(h.data, 0, (size_t)s.n * 3 * Vh * prec_cbytes(d.prec), stream()));
  if (h.norm) QOP_CHECK_HIP(hipMemsetAsync(h.norm, 0, (size_t)s.n * Vh * sizeof(float), stream()));
}

void axpy(DevVec &d, double a, const DevVec &x, const DevVec &y, int eo) {
  Cplx c[2] = {{a, 0}, {1, 0}};
  const DevVec *v[2] = {&x, &y};
  lincomb(d, eo, 2, c, v);
}

void caxpy(DevVec &d, Cplx a, const DevVec &x, int eo) {
  Cplx c[2] = {{1, 0}, a};
  const DevVec *v[2] = {&d, &x};
  lincomb(d, eo, 2, c, v);
}

void scal(DevVec &d, Cplx a, const DevVec &x, int eo) {
  const DevVec *v[1] = {&x};
  lincomb(d, eo, 1, &a, v);
}

}  // namespace qop
