// Batched multi-shift CG driver (msbatch.h) and the multi-mass entry point
// that refines every (source, mass) to its FP64 true residual.
#include "msbatch.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>

#include "handles.h"

namespace qop {

MultiShiftBatch::MultiShiftBatch(Links &L, int eo, int capacity, int sloppy_prec,
                                 const double *masses, int ns)
    : L_(L), lat_(L.lat), eo_(eo), cap_(capacity), sp_(sloppy_prec), ns_(ns),
      hp_(sloppy_prec == P32 && options().ms_link16 ? P32H : sloppy_prec) {
  try {
    QOP_REQUIRE(eo == 0 || eo == 1, "multi-shift: subset must be even or odd");
    QOP_REQUIRE(ns >= 1 && ns <= MS_MAXSHIFT, "multi-shift: 1..16 masses");
    QOP_REQUIRE(capacity >= 1 && capacity <= 64, "multi-shift: capacity 1..64");
    QOP_REQUIRE(sp_ == P32 || sp_ == P16, "multi-shift: sloppy precision must be FP32 or FP16");
    masses_.assign(masses, masses + ns);
    imin_ = 0;
    for (int i = 1; i < ns; i++)
      if (masses[i] * masses[i] < masses[imin_] * masses[imin_]) imin_ = i;
    nch_ = (int)((lat_->Vh + 63) / 64);
    devbatch_alloc(b64_, lat_, P64, cap_);
    devbatch_alloc(rs_, lat_, sp_, cap_);
    devbatch_alloc(ps_, lat_, sp_, cap_);
    devbatch_alloc(ts_, lat_, sp_, cap_);
    devbatch_alloc(pm_, lat_, sp_, cap_ * ns_);
    devbatch_alloc(xm_, lat_, P32, cap_ * ns_);
    devbatch_alloc(x64_, lat_, P64, cap_ * ns_);
    devbatch_alloc(t64_, lat_, P64, cap_);
    devbatch_alloc(ax64_, lat_, P64, cap_);
    dev_alloc(&ru_, cap_ * sizeof(int));
    QOP_CHECK_HIP(hipMemset(ru_, 0, cap_ * sizeof(int)));
    dev_alloc(&slot_, cap_ * sizeof(MsSlot));
    QOP_CHECK_HIP(hipHostMalloc(&slot_h_, cap_ * sizeof(MsSlot), hipHostMallocDefault));
    dev_alloc(&active_, 2 * cap_ * sizeof(int));
    sel_ = active_ + cap_;
    QOP_CHECK_HIP(hipHostMalloc(&active_h_, cap_ * sizeof(int), hipHostMallocDefault));
    dev_alloc(&partA_, 3 * (size_t)cap_ * nch_ * sizeof(double));
    partB_ = partA_ + (size_t)cap_ * nch_;
    partC_ = partB_ + (size_t)cap_ * nch_;
    dev_alloc(&coef_, 2 * (size_t)cap_ * sizeof(double));
    dev_alloc(&sums_, 2 * (size_t)cap_ * sizeof(double));
    dev_alloc(&zero_status_, sizeof(int));
    QOP_CHECK_HIP(hipMemset(zero_status_, 0, sizeof(int)));
    memset(slot_h_, 0, cap_ * sizeof(MsSlot));
    memset(active_h_, 0, cap_ * sizeof(int));
    QOP_CHECK_HIP(hipMemset(active_, 0, 2 * cap_ * sizeof(int)));
    const double m0 = masses[imin_];
    std::vector<double> hc(2 * cap_);
    for (int b = 0; b < cap_; b++) { hc[2 * b] = -1.0; hc[2 * b + 1] = m0 * m0; }
    QOP_CHECK_HIP(hipMemcpy(coef_, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice));
    L_.get(hp_);  // convert the links before any graph capture
    L_.get(P64);
    shape_ = tune_hop(L_, hp_, cap_, ps_, ts_, 1 - eo_);
    ru_shape_ = tune_hop(L_, P64, cap_, x64_, t64_, 1 - eo_);
    sync();
  } catch (const DeviceOutOfMemory &) {
    release();  // a failed allocation frees what was built
    throw;
  }
}

MultiShiftBatch::~MultiShiftBatch() { release(); }

void MultiShiftBatch::release() {
  sync();
  graph_.reset();
  for (DevBatch *b : {&b64_, &rs_, &ps_, &ts_, &pm_, &xm_, &x64_, &t64_, &ax64_}) devbatch_free(*b);
  (void)hipFree(ru_);
  (void)hipFree(slot_);
  (void)hipHostFree(slot_h_);
  (void)hipFree(active_);
  (void)hipHostFree(active_h_);
  (void)hipFree(partA_);
  (void)hipFree(coef_);
  (void)hipFree(sums_);
  (void)hipFree(zero_status_);
}

void MultiShiftBatch::load(int slot, const DevVec &src, double ms_tol, int max_iter) {
  QOP_REQUIRE(slot >= 0 && slot < cap_, "multi-shift: bad slot");
  hipStream_t st = stream();
  const long Vh = lat_->Vh;
  auto copy_view = [&](VecView dst, int pd, VecView s, int ps) {
    LincombArgs a;
    a.Vh = (int)Vh;
    a.out = dst;
    a.v[0] = s;
    a.cr[0] = 1.0;
    int pin[1] = {ps};
    launch_lincomb(a, pd, pin, 1, false, st);
  };
  copy_view(b64_.slot(slot), P64, src.half(eo_), P64);
  copy_view(rs_.slot(slot), sp_, src.half(eo_), P64);   // r = b (x = 0)
  copy_view(ps_.slot(slot), sp_, rs_.slot(slot), sp_);  // p = r
  for (int i = 0; i < ns_; i++) {
    const int k = i * cap_ + slot;
    QOP_CHECK_HIP(hipMemsetAsync(xm_.slot(k).data, 0, 3 * Vh * sizeof(float2), st));
    QOP_CHECK_HIP(hipMemsetAsync(x64_.slot(k).data, 0, 3 * Vh * sizeof(double2), st));
    if (i != imin_) copy_view(pm_.slot(k), sp_, rs_.slot(slot), sp_);
  }
  // |b|^2 (FP64), |r|^2 and |p|^2 of the stored sloppy vectors
  std::vector<int> sel(cap_, 0);
  sel[slot] = 1;
  QOP_CHECK_HIP(hipMemcpyAsync(sel_, sel.data(), cap_ * sizeof(int), hipMemcpyHostToDevice, st));
  DotArgs d;
  d.Vh = (int)Vh;
  d.n = cap_;
  d.ustride = 3 * Vh;
  d.nstride = Vh;
  d.active = sel_;
  d.u = b64_.slot(0);
  d.partial = partA_;
  launch_dot(d, P64, P64, 0, st);
  bs_launch_sum(partA_, nch_, cap_, sel_, sums_, st);
  d.u = ps_.slot(0);
  d.partial = partB_;  // |p|^2 rows for the first iteration
  launch_dot(d, sp_, sp_, 0, st);
  bs_launch_sum(partB_, nch_, cap_, sel_, sums_ + cap_, st);
  std::vector<double> h(2 * cap_);
  QOP_CHECK_HIP(hipMemcpyAsync(h.data(), sums_, 2 * cap_ * sizeof(double), hipMemcpyDeviceToHost, st));
  sync();
  MsSlot &s = slot_h_[slot];
  memset(&s, 0, sizeof s);
  s.ns = ns_;
  s.imin = imin_;
  s.max_iter = max_iter;
  s.b2 = h[slot];
  s.r2 = h[cap_ + slot];  // |r|^2 = |p|^2 of the stored sloppy r
  s.r2_max = s.r2;
  s.delta2 = options().delta * options().delta;
  const double m02 = masses_[imin_] * masses_[imin_];
  s.m2 = m02;
  for (int i = 0; i < ns_; i++) {
    s.shift[i] = masses_[i] * masses_[i] - m02;
    s.stop2[i] = ms_tol * ms_tol * s.b2;
    s.z[i] = s.zo[i] = s.zn[i] = 1;
    s.bo[i] = -1;
    s.a[i] = 0;
    s.b[i] = 0;
    s.live_next[i] = 1;
  }
  const bool zero = s.b2 == 0.0;
  if (zero) s.done = 1;
  active_h_[slot] = zero ? 0 : 1;
  QOP_CHECK_HIP(hipMemcpyAsync(slot_ + slot, &s, sizeof s, hipMemcpyHostToDevice, st));
  QOP_CHECK_HIP(hipMemcpyAsync(active_ + slot, active_h_ + slot, sizeof(int), hipMemcpyHostToDevice, st));
  sync();
}

void MultiShiftBatch::enqueue_iteration() {
  hipStream_t st = stream();
  LinkSet &ls = L_.get(hp_);
  const int opp = 1 - eo_;
  const long Vh = lat_->Vh;
  HopArgs a;
  a.g = lat_->g;
  a.parity = opp;
  a.links = ls.parity_data(lat_, opp);
  a.lscale = ls.parity_scale(lat_, opp);
  a.src = ps_.slot(0);
  a.dst = ts_.slot(0);
  a.sstride = a.dstride = a.xstride = 3 * Vh;
  a.nstride = Vh;
  a.n = cap_;
  a.active = active_;
  a.partial = partA_;
  shape_.apply(a);
  launch_hop_args(a, hp_, ls.nl, shape_.nb, shape_.block, RED_OUT2, false, st);
  ms_launch_scalarA(slot_, active_, partA_, partB_, nch_, cap_, st);
  // r -= beta_0 (m_0^2 p - D t), x_base += beta_0 p, |r|^2
  HopArgs b = a;
  b.parity = eo_;
  b.links = ls.parity_data(lat_, eo_);
  b.lscale = ls.parity_scale(lat_, eo_);
  b.src = ts_.slot(0);
  b.dst = rs_.slot(0);
  b.x = ps_.slot(0);
  b.xs = xm_.slot(imin_ * cap_);
  b.coef = coef_;
  b.cg_alpha = (const double *)((const char *)slot_ + offsetof(MsSlot, beta0));
  b.cg_stride = (int)(sizeof(MsSlot) / sizeof(double));
  b.partial = partC_;
  launch_hop_args(b, hp_, ls.nl, shape_.nb, shape_.block, RED_CG, true, st);
  ms_launch_scalarB(slot_, active_, partC_, nch_, cap_, ru_, st);
  ms_launch_update(slot_, active_, sp_, (int)Vh, rs_.slot(0), ps_.slot(0), pm_.slot(0), xm_.slot(0),
                   3 * Vh, 3 * Vh * cap_, cap_, ns_, partB_, st);
  enqueue_reliable_update();
}

// Masked by ru_: slots without a request skip every kernel (a masked hop group
// with no flagged slot reads nothing).
void MultiShiftBatch::enqueue_reliable_update() {
  hipStream_t st = stream();
  LinkSet &ls = L_.get(P64);
  const int opp = 1 - eo_;
  const long Vh = lat_->Vh;
  ms_launch_ru_fold(ru_, (int)Vh, x64_.slot(0), xm_.slot(0), 3 * Vh, cap_, ns_, st);
  HopArgs a;
  a.g = lat_->g;
  a.parity = opp;
  a.links = ls.parity_data(lat_, opp);
  a.lscale = ls.parity_scale(lat_, opp);
  a.src = x64_.slot(imin_ * cap_);
  a.dst = t64_.slot(0);
  a.sstride = a.dstride = a.xstride = 3 * Vh;
  a.nstride = Vh;
  a.n = cap_;
  a.active = ru_;
  ru_shape_.apply(a);
  launch_hop_args(a, P64, ls.nl, ru_shape_.nb, ru_shape_.block, RED_NONE, false, st);
  HopArgs b = a;
  b.parity = eo_;
  b.links = ls.parity_data(lat_, eo_);
  b.lscale = ls.parity_scale(lat_, eo_);
  b.src = t64_.slot(0);
  b.dst = ax64_.slot(0);
  b.x = x64_.slot(imin_ * cap_);
  b.coef = coef_;  // A_0 x = m_0^2 x - D t
  launch_hop_args(b, P64, ls.nl, ru_shape_.nb, ru_shape_.block, RED_NONE, true, st);
  ms_launch_ru_resid(ru_, sp_, (int)Vh, b64_.slot(0), ax64_.slot(0), rs_.slot(0), 3 * Vh, cap_,
                     partC_, st);
  ms_launch_ru_scalar(slot_, ru_, partC_, nch_, cap_, st);
}

void MultiShiftBatch::run() {
  hipStream_t st = stream();
  const int chunk = options().chunk > 0 ? options().chunk : 16;
  const bool use_graph = options().use_graph != 0;
  if (use_graph && !graph_) {
    graph_.begin(st);
    for (int c = 0; c < chunk; c++) enqueue_iteration();
    graph_.end(st);
  }
  while (true) {
    bool any = false;
    for (int i = 0; i < cap_; i++) any |= active_h_[i] != 0;
    if (!any) break;
    if (use_graph) graph_.launch(st);
    else for (int c = 0; c < chunk; c++) enqueue_iteration();
    QOP_CHECK_HIP(hipMemcpyAsync(active_h_, active_, cap_ * sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
  }
  QOP_CHECK_HIP(hipMemcpy(slot_h_, slot_, cap_ * sizeof(MsSlot), hipMemcpyDeviceToHost));
}

void MultiShiftBatch::solution(int slot, int k, DevVec &out) {
  QOP_REQUIRE(out.prec == P64, "multi-shift: solutions are FP64");
  LincombArgs a;
  a.Vh = (int)lat_->Vh;
  a.out = out.half(eo_);
  a.v[0] = x64_.slot(k * cap_ + slot);  // folded at the reliable updates
  a.v[1] = xm_.slot(k * cap_ + slot);   // increments since the last one
  a.cr[0] = 1.0;
  a.cr[1] = 1.0;
  int pin[2] = {P64, P32};
  launch_lincomb(a, P64, pin, 2, false, stream());
}

int MultiShiftBatch::updates(int slot) { return slot_h_[slot].nupdate; }

int MultiShiftBatch::iterations(int slot) { return slot_h_[slot].iter; }

}  // namespace qop

using namespace qop;

extern "C" {

// Multi-mass solves of nsrc sources over one shared mass list:
// out[i*nmass + j] = m_j A_j^-1 in[i] on the EVEN or ODD subset, each to FP64
// true relative residual sqrt(res[i*nmass+j]->rsqmin).  A multi-shift CG
// (FP32 vectors; FP16 with option qmx_ms_prec = 16) with reliable updates,
// batched over the sources, runs to relative residual
// ms_tol (0: the smallest requested tolerance); then every (source, mass) is
// checked in FP64 and, where needed, refined from its multi-shift solution by
// the reliable-update CG in sloppy precision `prec`, `capacity` systems at a
// time.  (An FP16 multi-shift stage stalls near 1e-3 and leaves most of the
// work to the refinement.)
int QOPX_D3_asqtad_invert_multi_batch(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *fla,
                                      QOP_invert_arg_t *inv_arg, QOP_resid_arg_t *res[],
                                      double masses[], int nmass, QOP_D3_ColorVector *out[],
                                      QOP_D3_ColorVector *in[], int nsrc, int capacity, int prec,
                                      double ms_tol, int *ms_iters) {
  Links &L = fla->L;
  const int eo = inv_arg->evenodd;
  QOP_REQUIRE(eo == QOP_EVEN || eo == QOP_ODD, "invert_multi_batch: subset must be EVEN or ODD");
  QOP_REQUIRE(prec == 16 || prec == 32, "invert_multi_batch: sloppy precision must be 16 or 32");
  QOP_REQUIRE(nmass >= 1 && nmass <= MS_MAXSHIFT && nsrc >= 0 && capacity >= 1, "invert_multi_batch: sizes");
  if (nsrc == 0) return 0;
  const int sp = prec == 16 ? P16 : P32;
  // with reliable updates the multi-shift stage can go to the requested
  // tolerance; the refinement then only finishes shifts whose solution drifted
  double tol_min = 1.0;
  for (int k = 0; k < nsrc * nmass; k++) tol_min = std::min(tol_min, std::sqrt(res[k]->rsqmin));
  if (ms_tol <= 0) ms_tol = tol_min;
  double t0 = wall_time();
  Lattice *lat = L.lat;
  const int n = nsrc * nmass;
  std::vector<DevVec> guess(n);
  // the multi-shift solutions are freed on every exit, also when a device
  // allocation fails
  struct FreeGuesses {
    std::vector<DevVec> &g;
    ~FreeGuesses() {
      for (DevVec &v : g) devvec_free(v);
    }
  } free_guesses{guess};
  long ms_total = 0, iters = 0;
  int ok = 0;
  const bool fit = guard_oom(__func__, info, [&] {
    // stage 1: multi-shift, sources in groups of up to `capacity`
    {
      const int cap = std::min(capacity, nsrc);
      MultiShiftBatch ms(L, eo, cap, options().ms_prec, masses, nmass);
      for (int s0 = 0; s0 < nsrc; s0 += cap) {
        const int cnt = std::min(cap, nsrc - s0);
        for (int b = 0; b < cnt; b++) ms.load(b, in[s0 + b]->v, ms_tol, inv_arg->max_iter);
        ms.run();
        for (int b = 0; b < cnt; b++) {
          if (ms_iters) ms_iters[s0 + b] = ms.iterations(b);
          ms_total += ms.iterations(b);
          for (int j = 0; j < nmass; j++) {
            DevVec &g = guess[(s0 + b) * nmass + j];
            devvec_alloc(g, lat, P64);
            ms.solution(b, j, g);
          }
        }
        sync();
      }
    }
    // stage 2: FP64 reliable-update refinement of every (source, mass)
    const int cap = std::min(capacity, n);
    BatchSolver bs(L, eo, cap, sp, options().delta);
    std::vector<int> slot_of(cap, -1);
    int next = 0, finished = 0;
    std::vector<BatchSolver::LoadReq> reqs;
    while (finished < n) {
      reqs.clear();
      for (int s = 0; s < cap && next < n; s++) {
        if (!bs.slot_free(s)) continue;
        const int src = next / nmass, j = next % nmass;
        reqs.push_back({s, &in[src]->v, &guess[next], masses[j], std::sqrt(res[next]->rsqmin),
                        inv_arg->max_iter});
        slot_of[s] = next++;
      }
      bs.load_many(reqs.data(), (int)reqs.size());
      std::vector<int> done;
      bs.step(done);
      for (int d : done) {
        const int k = slot_of[d];
        SlotResult r;
        bs.result(d, out[k]->v, true, &r);
        res[k]->final_iter = r.iter;
        res[k]->final_rsq = r.true_rsq;
        res[k]->final_rel = 0;
        res[k]->final_restart = r.nupdate;
        ok += r.reason == 1;
        iters += r.iter;
        bs.release(d);
        devvec_free(guess[k]);
        finished++;
      }
    }
  });
  if (!fit) return -1;
  if (info) {
    info->final_sec = wall_time() - t0;
    const double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl), nflopm = 0.5 * 10 * 3;
    info->final_flop = ((nflop + nflopm * (nmass - 1)) * ms_total + nflop * iters) * (double)lat->V;
    info->status = QOP_SUCCESS;
  }
  return ok;
}

}  // extern "C"
