// The reference solver algorithms on the GPU, with the same control flow and
// stopping rules as QOPQDP:
//   invert_cgms  -- lib/invert_cg_p.c:130-402 (scalar recurrences on device)
//   invert_cg    -- lib/invert_cg_p.c:1-128
//   solve_multi  -- lib/asqtad_invert_p.c:246-647 (with refine)
//   invert       -- lib/asqtad_invert_p.c:134-144, 649-851
//   invert_multi -- lib/asqtad_invert_p.c:853-1059
// The vectors of a solve have the precision of the caller's vectors: FP64 for
// the QOP_D3 entry points, FP32 for QOP_F3 (lib/pcstubs/asqtad_invert_f3.c
// instantiates the same drivers in single precision); the scalars are FP64
// throughout.  The FP32 branch of an FP64 solve_multi (mixed_rsq) runs the
// same multi-shift CG on FP32 vectors and FP32 links, as the reference does.
#include <memory>
#include <chrono>
#include <cmath>
#include <cstring>

#include "solver.h"
#include "eigcg.h"

namespace qop {

double wall_time() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

Options &options() {
  static Options o;
  return o;
}

// <p, A p>, out = A p = m^2 p - D_eo D_oe p on subset eo; tmp gets D_oe p.
double ddagd_norm2(Links &L, int prec, double mass, DevVec &out, const DevVec &in,
                   int eo, DevVec &tmp) {
  QOP_REQUIRE(eo == 0 || eo == 1, "ddagd: subset must be even or odd");
  double *res = red_result(2);
  // red_result is shared; use two distinct slots via the finalize target
  hop(L, prec, tmp, 1 - eo, in, 1.0, 0.0, nullptr, RED_OUT2, res, stream());
  hop(L, prec, out, eo, tmp, -1.0, mass * mass, &in, RED_X2, res + 1, stream());
  double h[2];
  QOP_CHECK_HIP(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  sync();
  return h[0] + mass * mass * h[1];
}

// ---------------------------------------------------------------------------
// multi-shift CG
namespace {

struct CgmsBuffers {
  CgmsDev d{};
  CgmsState *host = nullptr;
  int nch = 0;
  CgmsBuffers(Lattice *lat, int ns) {
    nch = (int)((lat->Vh + 63) / 64);
    try {
      dev_alloc(&d.st, sizeof(CgmsState));
      dev_alloc(&d.partA, 4 * (size_t)nch * sizeof(double));
      d.partB = d.partA + nch;
      d.partC = d.partB + nch;
      d.partD = d.partC + nch;
      dev_alloc(&d.xs, 2 * (size_t)ns * sizeof(VecView));
      d.pms = d.xs + ns;
    } catch (const DeviceOutOfMemory &) {
      release();
      throw;
    }
    QOP_CHECK_HIP(hipHostMalloc(&host, sizeof(CgmsState), hipHostMallocDefault));
  }
  ~CgmsBuffers() { release(); }
  void release() {
    if (d.st) (void)hipFree(d.st);
    if (d.partA) (void)hipFree(d.partA);
    if (d.xs) (void)hipFree(d.xs);
    if (host) (void)hipHostFree(host);
    d.st = nullptr;
    d.partA = nullptr;
    d.xs = nullptr;
    host = nullptr;
  }
  void pull(hipStream_t st) {
    QOP_CHECK_HIP(hipMemcpyAsync(host, d.st, sizeof(CgmsState), hipMemcpyDeviceToHost, st));
    QOP_CHECK_HIP(hipStreamSynchronize(st));
  }
  void push(hipStream_t st) {
    QOP_CHECK_HIP(hipMemcpyAsync(d.st, host, sizeof(CgmsState), hipMemcpyHostToDevice, st));
  }
};

}  // namespace

void invert_cgms(Links &L, int prec, double mass, int eo, InvArg *ia, ResArg **ra,
                 const double *shifts, int ns, DevVec **out, const DevVec &in) {
  Lattice *lat = L.lat;
  hipStream_t st = stream();
  QOP_REQUIRE(ns >= 1 && ns <= CGMS_MAXSHIFT, "invert_cgms: 1..64 shifts");
  QOP_REQUIRE(eo == 0 || eo == 1, "invert_cgms: subset must be even or odd");
  const int opp = 1 - eo;
  double insq = norm2(in, eo);
  if (insq == 0.) {  // :152-162
    for (int i = 0; i < ns; i++) {
      zero(*out[i], eo);
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
  const double rsqstop = ra[irsq]->rsqmin * insq;
  const double relstop = ra[irel]->relmin;

  TmpVec r(lat, prec), Mp(lat, prec), tmp(lat, prec), p(lat, prec);
  TmpVec y(lat, prec), Ay(lat, prec), ry(lat, prec);
  std::vector<std::unique_ptr<TmpVec>> pm(ns);
  copy(r.v, in, eo);
  for (int i = 0; i < ns; i++) {
    if (i != imin) pm[i].reset(new TmpVec(lat, prec));
    DevVec &pv = i == imin ? p.v : pm[i]->v;
    zero(*out[i], eo);
    copy(pv, r.v, eo);
  }
  copy(ry.v, in, eo);

  CgmsBuffers B(lat, ns);
  std::vector<VecView> hx(ns), hp(ns);
  for (int i = 0; i < ns; i++) {
    hx[i] = out[i]->half(eo);
    hp[i] = (i == imin ? p.v : pm[i]->v).half(eo);
  }
  QOP_CHECK_HIP(hipMemcpyAsync(B.d.xs, hx.data(), ns * sizeof(VecView), hipMemcpyHostToDevice, st));
  QOP_CHECK_HIP(hipMemcpyAsync(B.d.pms, hp.data(), ns * sizeof(VecView), hipMemcpyHostToDevice, st));
  CgmsState &S = *B.host;
  memset(&S, 0, sizeof S);
  S.status = 0;
  S.iteration = 0;
  S.ns = ns; S.imin = imin; S.irsq = irsq; S.irel = irel;
  S.restart = ia->restart; S.max_iter = ia->max_iter;
  S.nextRsqCheck = 100; S.nextRelCheck = 100;
  S.rsq = insq; S.insq = insq; S.rsqstop = rsqstop; S.relstop = relstop;
  S.trueRsqFac = 1; S.trueRelFac = 1; S.relnorm2 = 1;
  S.m2 = mass * mass;
  for (int i = 0; i < ns; i++) {
    S.shifts[i] = shifts[i];
    S.zo[i] = S.z[i] = 1; S.zn[i] = 1;
    S.bo[i] = -1; S.a[i] = 0; S.b[i] = 0;
  }
  B.push(st);

  const int Vh = (int)lat->Vh;
  const int nl = L.nl;
  const int with_rel = relstop > 0;
  // link arguments of both hops, resolved (and any link conversion done)
  // before the iterations are captured into a graph
  HopArgs l_opp, l_eo;
  set_hop_links(l_opp, L, prec, opp);
  set_hop_links(l_eo, L, prec, eo);
  auto enqueue_iteration = [&]() {
    // linop: tmp = D p (|tmp|^2), Mp = m^2 p - D tmp (|p|^2)
    HopArgs a1 = l_opp;
    a1.g = lat->g; a1.parity = opp;
    a1.src = p.v.half(eo); a1.dst = tmp.v.half(opp);
    a1.partial = B.d.partA; a1.status = &B.d.st->status;
    launch_hop_args(a1, prec, nl, 1, 256, RED_OUT2, false, st);
    HopArgs a2 = l_eo;
    a2.g = lat->g; a2.parity = eo;
    a2.src = tmp.v.half(opp); a2.dst = Mp.v.half(eo);
    a2.x = p.v.half(eo); a2.alpha = -1.0; a2.beta = mass * mass;
    a2.partial = B.d.partB; a2.status = &B.d.st->status;
    launch_hop_args(a2, prec, nl, 1, 256, RED_X2, true, st);
    cgms_launch_scalarA(B.d, B.nch, st);
    cgms_launch_update(B.d, prec, Vh, r.v.half(eo), Mp.v.half(eo), ns, st);
    if (with_rel) cgms_launch_relnorm(B.d, prec, Vh, r.v.half(eo), out[irel]->half(eo), st);
    cgms_launch_scalarB(B.d, B.nch, with_rel, Vh, st);
    cgms_launch_pupdate(B.d, prec, Vh, p.v.half(eo), r.v.half(eo), ns, st);
  };

  const int chunk = options().chunk > 0 ? options().chunk : 16;
  CapturedGraph gexec;
  if (options().use_graph) {
    gexec.begin(st);
    for (int c = 0; c < chunk; c++) enqueue_iteration();
    gexec.end(st);
  }

  double yAy = 0;
  auto linop = [&](DevVec &o, const DevVec &i) { return ddagd_norm2(L, prec, mass, o, i, eo, tmp.v); };
  DevVec &xmin = *out[imin];
  while (true) {
    if (gexec) gexec.launch(st);
    else for (int c = 0; c < chunk; c++) enqueue_iteration();
    B.pull(st);
    if (S.status == 2) break;
    if (S.status == 0) continue;
    // status 1: fold the base solution into y and check the true residual
    // (:277-346), on the host with device vector operations.
    bool stop = false;
    int &it = S.iteration;
    if (S.nextRsqCheck <= it || (relstop > 0 && S.nextRelCheck <= it)) {
      if (yAy > 0) {
        Cplx yAx = cdot(Ay.v, xmin, eo);
        caxpy(xmin, Cplx{-yAx.re / yAy, -yAx.im / yAy}, y.v, eo);
      }
      double xAx = linop(Mp.v, xmin);
      it++;
      if (xAx <= 0) { stop = true; }
      if (!stop) {
        Cplx xr = cdot(xmin, ry.v, eo);
        caxpy(y.v, Cplx{xr.re / xAx, xr.im / xAx}, xmin, eo);
        yAy = linop(Ay.v, y.v);
        it++;
        axpy(ry.v, -1.0, Ay.v, in, eo);
        zero(xmin, eo);
      }
    }
    if (!stop && S.nextRsqCheck <= it) {
      double tr2;
      if (irsq == imin) {
        tr2 = norm2(ry.v, eo);
      } else {
        double xAx = linop(Mp.v, *out[irsq]);
        it++;
        if (xAx <= 0) stop = true;
        axpy(Mp.v, -1.0, in, Mp.v, eo);
        tr2 = norm2(Mp.v, eo);
      }
      if (!stop) {
        if (options().verbose) printf("CGMS: %i: rsq: %g  tr2: %g\n", it, S.rsq, tr2);
        if (tr2 <= rsqstop) { S.rsq = tr2; stop = true; }
        else if (tr2 > 2 * S.rsq) stop = true;
        else { S.trueRsqFac = tr2 / S.rsq; S.nextRsqCheck += 100; }
      }
    }
    if (!stop && relstop > 0 && S.nextRelCheck <= it) {
      double xAx = linop(Mp.v, *out[irel]);
      it++;
      if (xAx <= 0) stop = true;
      else {
        axpy(Mp.v, -1.0, in, Mp.v, eo);
        double rel2 = relnorm2(Mp.v, *out[irel], eo);
        if (rel2 <= relstop) { S.relnorm2 = rel2; stop = true; }
        else { S.trueRelFac = rel2 / S.relnorm2; S.nextRelCheck += 100; S.relnorm2 = rel2; }
      }
    }
    if (stop) { S.status = 2; break; }
    // alpha and shift-over (:349-376), then p updates
    S.a[imin] = S.rsq / S.oldrsq;
    for (int i = 0; i < ns; i++) {
      if (i == imin) continue;
      double c2 = S.z[i] * S.b[imin];
      S.a[i] = c2 != 0.0 ? S.a[imin] * S.zn[i] * S.b[i] / c2 : 0;
    }
    for (int i = 0; i < ns; i++) { S.bo[i] = S.b[i]; S.zo[i] = S.z[i]; S.z[i] = S.zn[i]; }
    S.status = 0;
    B.push(st);
    cgms_launch_pupdate(B.d, prec, Vh, p.v.half(eo), r.v.half(eo), ns, st);
  }
  gexec.reset();
  copy(xmin, y.v, eo);  // :378
  sync();
  for (int i = 0; i < ns; i++) {
    double s = S.zn[i] * S.zn[i];
    ra[i]->final_rsq = s * S.trueRsqFac * S.rsq / insq;
    ra[i]->final_rel = s * S.trueRelFac * S.relnorm2;
    ra[i]->final_iter = S.iteration;
    ra[i]->final_restart = 0;
  }
}

// ---------------------------------------------------------------------------
// regular CG (lib/invert_cg_p.c:1-128), linop = ddagd
void invert_cg(Links &L, int prec, double mass, int eo, InvArg *ia, ResArg *ra,
               DevVec &out, const DevVec &in) {
  Lattice *lat = L.lat;
  double a, b, rsq, oldrsq, pkp, relnorm = 1.0, insq, rsqstop;
  int iteration = 0, total = 0, nrestart = -1;
  int restart = ia->restart, max_iter = ia->max_iter, max_restarts = ia->max_restarts;
  if (max_restarts < 0) max_restarts = 5;
  ra->final_rsq = 0; ra->final_rel = 0; ra->final_iter = 0; ra->final_restart = 0;
  insq = norm2(in, eo);
  if (insq == 0.) { zero(out, eo); return; }
  TmpVec r(lat, prec), Mp(lat, prec), p(lat, prec), tmp(lat, prec);
  auto linop = [&](DevVec &o, const DevVec &i) { ddagd_norm2(L, prec, mass, o, i, eo, tmp.v); };
  rsqstop = ra->rsqmin * insq;
  rsq = 0; oldrsq = rsq;
  while (true) {
    if (total == 0 || iteration >= restart || total >= max_iter || rsq < rsqstop ||
        relnorm < ra->relmin) {
      if (total >= max_iter || nrestart >= max_restarts) break;
      nrestart++;
      copy(p.v, out, eo);
      linop(Mp.v, p.v);
      iteration = 1;
      total++;
      axpy(r.v, -1.0, Mp.v, in, eo);
      rsq = norm2(r.v, eo);
      if (ra->relmin > 0) relnorm = relnorm2(r.v, out, eo);
      if (rsq < rsqstop || relnorm < ra->relmin || total >= max_iter) break;
      copy(p.v, r.v, eo);
    } else {
      b = rsq / oldrsq;
      axpy(p.v, b, p.v, r.v, eo);
    }
    oldrsq = rsq;
    linop(Mp.v, p.v);
    iteration++;
    total++;
    pkp = redot(p.v, Mp.v, eo);
    if (pkp <= 0) break;
    a = rsq / pkp;
    axpy(out, a, p.v, out, eo);
    axpy(r.v, -a, Mp.v, r.v, eo);
    rsq = norm2(r.v, eo);
    if (ra->relmin > 0) relnorm = relnorm2(r.v, out, eo);
  }
  ra->final_rsq = rsq / insq;
  ra->final_rel = relnorm;
  ra->final_iter = total;
  ra->final_restart = nrestart;
}

// ---------------------------------------------------------------------------
// refine (lib/asqtad_invert_p.c:246-330).  Aout[i] keeps A-side vectors on the
// opposite parity (the reference stores them in r[i] on that parity).
static void refine(Links &L, DevVec **out, DevVec *const *in, DevVec **r, DevVec **Aout,
                   double *r2, const double *m2, Cplx *scale, int no, int ineo,
                   DevVec **x, DevVec **Ax, int nx, const int *oi, const int *xi, int ni,
                   DevVec &tv) {
  int opeo = oppsub(ineo);
  std::vector<double> x2(nx + 1), Ax2(nx + 1);
  for (int i = 0; i < nx; i++) {
    dslash(L, 0, *Ax[i], *x[i], opeo, ineo);
    x2[i] = norm2(*x[i], ineo);
    Ax2[i] = norm2(*Ax[i], opeo);
  }
  for (int i = 0; i < ni; i++) {
    int ko = oi[i];
    int kx = nx > 0 ? xi[i] : 0;
    Cplx sx{0, 0}, z{0, 0}, Az{0, 0};
    double d = 0;
    dslash(L, 0, *Aout[ko], *out[ko], opeo, ineo);
    if (nx > 0) {
      z = cdot(*x[kx], *out[ko], ineo);
      Az = cdot(*Ax[kx], *Aout[ko], opeo);
      d = Ax2[kx] + m2[ko] * x2[kx];
    }
    if (d != 0) {
      double di = 1 / d, dim2 = di * m2[ko];
      Cplx s{dim2 * z.re + di * Az.re, dim2 * z.im + di * Az.im};
      caxpy(*out[ko], Cplx{-s.re, -s.im}, *x[kx], ineo);
      caxpy(*Aout[ko], Cplx{-s.re, -s.im}, *Ax[kx], opeo);
      z = cdot(*x[kx], *in[ko], ineo);
      sx = Cplx{di * z.re, di * z.im};
    }
    z = cdot(*out[ko], *in[ko], ineo);
    double o2 = norm2(*out[ko], ineo);
    double Ao2 = norm2(*Aout[ko], opeo);
    d = Ao2 + m2[kx] * o2;  // the reference indexes m2 with kx here (:299)
    if (d != 0) {
      double di = 1 / d;
      Cplx s{di * z.re, di * z.im};
      scale[ko] = s;
      if (nx > 0) {
        double den = s.re * s.re + s.im * s.im;
        Cplx q{(sx.re * s.re + sx.im * s.im) / den, (sx.im * s.re - sx.re * s.im) / den};
        caxpy(*out[ko], q, *x[kx], ineo);
        caxpy(*Aout[ko], q, *Ax[kx], opeo);
      }
    } else {
      scale[ko] = sx;
      if (nx > 0) {
        copy(*out[ko], *x[kx], ineo);
        copy(*Aout[ko], *Ax[kx], opeo);
      }
    }
  }
  for (int i = 0; i < no; i++) {  // r <- in - scale A out
    dslash(L, 0, tv, *Aout[i], ineo, opeo);
    axpy(tv, -m2[i], *out[i], tv, ineo);
    copy(*r[i], *in[i], ineo);
    caxpy(*r[i], scale[i], tv, ineo);
    r2[i] = norm2(*r[i], ineo);
  }
}

void solve_multi(Links &L, InvArg *ia, ResArg **ra, const double *masses, DevVec **out,
                 DevVec *const *in, int nsolve, double *final_flop, double *final_sec) {
  Lattice *lat = L.lat;
  const int ineo = ia->evenodd;
  const int bp = in[0]->prec;  // vector precision of this solve
  std::vector<double> in2(nsolve), r2(nsolve), r2stop(nsolve), mass2(nsolve), m2s(nsolve);
  std::vector<Cplx> scale(nsolve);
  std::vector<int> oi(nsolve), xi(nsolve);
  int nm = 0, iter = 0, nrestart = -1;
  const int max_iter = ia->max_iter, max_restarts = ia->max_restarts;
  const double mixedrsq = ia->mixed_rsq;
  double dtime = 0;
  std::vector<std::unique_ptr<TmpVec>> rv(nsolve), av(nsolve);
  TmpVec tv(lat, bp);
  for (int i = 0; i < nsolve; i++) {
    in2[i] = norm2(*in[i], ineo);
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
    rv[i].reset(new TmpVec(lat, bp));
    av[i].reset(new TmpVec(lat, bp));
  }
  for (int i = 0; i < nsolve; i++) {
    int j = 0;
    while (j < nm - 1 && m2s[j] != mass2[i]) j++;
    xi[i] = j;
  }
  ResArg resarg0{0, 0, 0, 0, 0, 0};
  std::vector<ResArg *> xresarg(nm);
  std::vector<std::unique_ptr<TmpVec>> xv(nm), axv(nm);
  double m20 = m2s[0];
  double gl_mass = std::sqrt(m20);
  for (int i = 0; i < nm; i++) {
    m2s[i] -= m20;
    if (bp == P32) m2s[i] = (float)m2s[i];  // QLA_Real shifts of the F build (:356)
    xv[i].reset(new TmpVec(lat, bp));
    axv[i].reset(new TmpVec(lat, bp));
  }
  std::vector<DevVec *> r(nsolve), Aout(nsolve), x(nm), Ax(nm);
  for (int i = 0; i < nsolve; i++) { r[i] = &rv[i]->v; Aout[i] = &av[i]->v; }
  for (int i = 0; i < nm; i++) { x[i] = &xv[i]->v; Ax[i] = &axv[i]->v; }
  refine(L, out, in, r.data(), Aout.data(), r2.data(), mass2.data(), scale.data(), nsolve,
         ineo, nullptr, nullptr, 0, oi.data(), nullptr, nsolve, tv.v);
  while (true) {
    for (int i = 0; i < nsolve; i++)
      ra[i]->final_rel = ra[i]->relmin > 0 ? relnorm2(*r[i], *out[i], ineo) : 1;
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
    if (options().verbose)
      printf("SOLVE: its: %i worst: %i mass: %g rsq: %g/%g\n", iter, imax, masses[imax],
             r2[imax] / in2[imax], ra[imax]->rsqmin);
    if (r2max <= r2stopmax || nrestart >= max_restarts || iter >= max_iter) break;
    for (int i = 0; i < nm; i++) {
      xresarg[i] = i == xi[imax] ? ra[imax] : &resarg0;
      zero(*x[i], 2);
    }
    double rsqminold = ra[imax]->rsqmin;
    ra[imax]->rsqmin = r2stopmax / r2max;
    if (ra[imax]->rsqmin >= 1) ra[imax]->rsqmin = in2max * r2stopmax / r2max;
    ra[imax]->rsqmin *= rsqminfac;
    ia->max_iter = max_iter - iter;
    // the FP32 inner branch exists in the double-precision build only (:510)
    const bool mixed = bp == P64 && r2max > mixedrsq * in2max;
    double t0 = wall_time();
    if (mixed) {
      // FP32 inner multi-shift CG (:510-538)
      if (mixedrsq > rsqminold) ra[imax]->rsqmin = mixedrsq * in2max / r2max;
      std::vector<double> m2sf(nm);
      for (int i = 0; i < nm; i++) m2sf[i] = (float)m2s[i];
      TmpVec rf(lat, P32);
      copy(rf.v, *r[imax], ineo);
      std::vector<std::unique_ptr<TmpVec>> xf(nm);
      std::vector<DevVec *> xfp(nm);
      for (int i = 0; i < nm; i++) { xf[i].reset(new TmpVec(lat, P32)); xfp[i] = &xf[i]->v; }
      invert_cgms(L, P32, gl_mass, ineo, ia, xresarg.data(), m2sf.data(), nm, xfp.data(), rf.v);
      for (int i = 0; i < nm; i++) { copy(*x[i], xf[i]->v, ineo); sync(); xf[i].reset(); }
    } else {
      invert_cgms(L, bp, gl_mass, ineo, ia, xresarg.data(), m2s.data(), nm, x.data(), *r[imax]);
    }
    dtime += wall_time() - t0;
    iter += ra[imax]->final_iter;
    nrestart++;
    ra[imax]->rsqmin = rsqminold;
    if (usingrel) {
      for (int i = 0; i < nsolve; i++) {
        int ko = oi[i], kx = xi[i];
        double s2 = scale[ko].re * scale[ko].re + scale[ko].im * scale[ko].im;
        if (s2 == 0.0) copy(*out[ko], *x[kx], ineo);
        else caxpy(*out[ko], Cplx{scale[ko].re / s2, -scale[ko].im / s2}, *x[kx], ineo);
      }
      refine(L, out, in, r.data(), Aout.data(), r2.data(), mass2.data(), scale.data(), nsolve,
             ineo, nullptr, nullptr, 0, oi.data(), nullptr, nsolve, tv.v);
    } else {
      refine(L, out, in, r.data(), Aout.data(), r2.data(), mass2.data(), scale.data(), nsolve,
             ineo, x.data(), Ax.data(), nm, oi.data(), xi.data(), nsolve, tv.v);
    }
  }
  ia->max_iter = max_iter;
  for (int i = 0; i < nsolve; i++) {
    Cplx s{masses[i] * scale[i].re, masses[i] * scale[i].im};
    scal(tv.v, s, *out[i], ineo);
    copy(*out[i], tv.v, ineo);
    ra[i]->final_iter = iter;
    ra[i]->final_rsq = in2[i] > 0 ? r2[i] / in2[i] : 0;
    ra[i]->final_restart = nrestart > 0 ? nrestart : 0;
  }
  sync();
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl), nflopm = 0.5 * 10 * 3;
  if (final_flop) *final_flop = (nflop + nflopm * (nm - 1)) * iter * (double)lat->V;
  if (final_sec) *final_sec = dtime;
}

// ---------------------------------------------------------------------------
void invert(Links &L, InvArg *ia, ResArg *ra, double mass, DevVec &out, const DevVec &in,
            double *final_flop, double *final_sec) {
  if (ia->evenodd != 2) {
    DevVec *o[1] = {&out};
    DevVec *i[1] = {const_cast<DevVec *>(&in)};
    ResArg *r[1] = {ra};
    solve_multi(L, ia, r, &mass, o, i, 1, final_flop, final_sec);
    return;
  }
  Lattice *lat = L.lat;
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl);
  const int ineo = ia->evenodd;
  const int bp = in.prec;
  int iter = 0, nrestart = -1;
  const int max_iter_old = ia->max_iter, max_restarts_old = ia->max_restarts;
  int max_restarts = ia->max_restarts;
  if (max_restarts <= 0) max_restarts = 5;
  double insqe = 0, insqo = 0;
  if (ineo != 1) insqe = norm2(in, 0);
  if (ineo != 0) insqo = norm2(in, 1);
  double insq = insqe + insqo;
  const int cgeo = insqe >= insqo ? 0 : 1;
  const int oe = oppsub(cgeo);
  TmpVec qin(lat, bp), qout(lat, bp), cgr(lat, bp), tmp2(lat, bp), gtmp(lat, bp);
  copy(tmp2.v, in, ineo);
  if (ineo != 2) zero(tmp2.v, oppsub(ineo));
  // project: qin = mass*tmp2 - D tmp2 on cgeo (:79-89)
  dslash(L, mass, qin.v, tmp2.v, cgeo, oe);
  {
    Cplx c[2] = {{mass, 0}, {-1.0, 0}};
    const DevVec *v[2] = {&tmp2.v, &qin.v};
    lincomb(qin.v, cgeo, 2, c, v);
  }
  if (ineo == cgeo) zero(qin.v, oe);
  else copy(qin.v, in, oe);
  double pinsq = norm2(qin.v, cgeo);
  double rsqfac = mass * mass * insq / pinsq;
  zero(qout.v, 2);
  copy(qout.v, out, cgeo);
  double rsqstop = insq * ra->rsqmin, rsq = 0, relnorm = 1.;
  double rsqminold = ra->rsqmin, relminold = ra->relmin;
  ra->rsqmin = 0.99 * rsqminold * rsqfac;
  ra->relmin = 0.99 * relminold;
  ia->max_restarts = 0;
  double dtime = 0;
  do {
    double t0 = wall_time();
    // eigCG keeps its deflation space in FP64; F3 solves run the plain CG
    if (options().cgtype == 1 && bp == P64) {
      // eigCG inner solves (asqtad_invert_p.c:683-687, 767-769); the sizes
      // are taken from the options at every call, as the reference does
      if (!L.eig) L.eig = new EigcgState;
      L.eig->numax = options().eigcg_numax;
      L.eig->m = options().eigcg_m;
      L.eig->nev = options().eigcg_nev;
      invert_eigcg(L, mass, cgeo, ia, ra, qout.v, qin.v, *L.eig);
    } else {
      invert_cg(L, bp, mass, cgeo, ia, ra, qout.v, qin.v);
    }
    sync();
    dtime += wall_time() - t0;
    iter += ra->final_iter;
    // reconstruct: qout_o = (qin_o - D qout_e)/mass (:91-100)
    copy(gtmp.v, qout.v, cgeo);
    dslash(L, mass, qout.v, gtmp.v, oe, cgeo);
    axpy(qout.v, -1.0, qout.v, qin.v, oe);
    scal(qout.v, Cplx{1.0 / mass, 0}, qout.v, oe);
    // full residual (:797-799)
    dslash(L, mass, cgr.v, qout.v, ineo, 2);
    axpy(cgr.v, -1.0, in, cgr.v, ineo);
    rsq = norm2(cgr.v, ineo);
    ra->rsqmin *= 0.99;
    if (ra->relmin > 0) {
      relnorm = relnorm2(cgr.v, qout.v, ineo);
      ra->relmin = 0.99 * ra->final_rel / relnorm * relminold;
    }
    ia->max_iter = max_iter_old - iter;
  } while (++nrestart < max_restarts && ia->max_iter > 0 && rsq > rsqstop &&
           relnorm > ra->relmin);
  copy(out, qout.v, ineo);
  sync();
  ia->max_iter = max_iter_old;
  ia->max_restarts = max_restarts_old;
  ra->rsqmin = rsqminold;
  ra->relmin = relminold;
  ra->final_iter = iter;
  ra->final_rsq = insq == 0. ? 0 : rsq / insq;
  ra->final_rel = relnorm;
  ra->final_restart = nrestart;
  if (final_flop) *final_flop = nflop * iter * (double)lat->V;
  if (final_sec) *final_sec = dtime;
}

// ---------------------------------------------------------------------------
void invert_multi(Links &L, InvArg *ia, ResArg **ra, const double *masses, const int *nmass,
                  DevVec **out, DevVec *const *in, int nsrc, double *final_flop,
                  double *final_sec) {
  Lattice *lat = L.lat;
  int nsolve = 0;
  for (int i = 0; i < nsrc; i++) nsolve += nmass[i];
  if (ia->evenodd != 2) {
    std::vector<DevVec *> inp(nsolve);
    int k = 0;
    for (int i = 0; i < nsrc; i++)
      for (int j = 0; j < nmass[i]; j++) inp[k++] = in[i];
    solve_multi(L, ia, ra, masses, out, inp.data(), nsolve, final_flop, final_sec);
    return;
  }
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl), nflopm = 0.5 * 10 * 3;
  const int bp = in[0]->prec;
  const int cgeo = 0;  // EVENODD "not supported yet": uses EVEN (:938-941)
  fprintf(stderr, "warning: QOP_asqtad_invert_multi on ALL subset not supported yet!\n");
  double flop = 0;
  double t0 = wall_time();
  if (nsrc == 2 && nmass[0] == 1 && nmass[1] == 1 && masses[0] != masses[1]) {
    double shifts[2];
    int imin = 0;
    for (int i = 0; i < 2; i++) {
      shifts[i] = masses[i] * masses[i];
      if (shifts[i] < shifts[imin]) imin = i;
    }
    double gm = masses[imin];
    for (int i = 0; i < 2; i++) shifts[i] -= gm * gm;
    double st = 1 / (shifts[1] - shifts[0]);
    TmpVec x0(lat, bp), src(lat, bp), cgp(lat, bp), tmp(lat, bp);
    axpy(x0.v, -1.0, *in[0], *in[1], cgeo);
    scal(x0.v, Cplx{st, 0}, x0.v, cgeo);
    copy(cgp.v, x0.v, cgeo);
    ddagd_norm2(L, bp, gm, src.v, cgp.v, cgeo, tmp.v);
    axpy(src.v, -1.0, src.v, *in[imin], cgeo);
    double rsq1 = norm2(src.v, cgeo), rsqsave[2];
    for (int i = 0; i < 2; i++) {
      double rsq0 = norm2(*in[i], cgeo);
      rsqsave[i] = ra[i]->rsqmin;
      ra[i]->rsqmin *= rsq0 / rsq1;
    }
    invert_cgms(L, bp, gm, cgeo, ia, ra, shifts, 2, out, src.v);
    flop += (nflop + nflopm) * ra[imin]->final_iter * (double)lat->V;
    for (int i = 0; i < 2; i++) {
      ra[i]->rsqmin = rsqsave[i];
      axpy(*out[i], 1.0, x0.v, *out[i], cgeo);
    }
    sync();
  } else {
    int k0 = 0;
    for (int i = 0; i < nsrc; i++) {
      int nm = nmass[i], jmin = 0;
      std::vector<double> shifts(nm);
      for (int j = 0; j < nm; j++) {
        shifts[j] = masses[k0 + j] * masses[k0 + j];
        if (shifts[j] < shifts[jmin]) jmin = j;
      }
      double gm = masses[k0 + jmin];
      for (int j = 0; j < nm; j++) shifts[j] -= gm * gm;
      invert_cgms(L, bp, gm, cgeo, ia, ra + k0, shifts.data(), nm, out + k0, *in[i]);
      flop += (nflop + nflopm * (nm - 1)) * ra[k0]->final_iter * (double)lat->V;
      k0 += nm;
    }
  }
  for (int k = 0; k < nsolve; k++) scal(*out[k], Cplx{masses[k], 0}, *out[k], cgeo);
  sync();
  if (final_flop) *final_flop = flop;
  if (final_sec) *final_sec = wall_time() - t0;
}

}  // namespace qop
