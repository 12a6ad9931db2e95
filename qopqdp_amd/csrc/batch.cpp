// Batched mixed-precision CG driver (see batch.h) and the hop autotuner.
#include "batch.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace qop {

// ---------------------------------------------------------------------------
// autotuner: time the multi-RHS hop over launch shapes and keep the best:
// first RHS per thread x block size, then remap / link loads / block order /
// site tiling for the best of those.
static std::mutex g_tune_mu;  // solvers may be built from several host threads
static bool g_override = false;
static HopShape g_override_shape;

// A forced launch shape (tests of shape independence); invalid for a given
// lattice or precision it is ignored there.
void set_hop_shape_override(const HopShape *s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_override = s != nullptr;
  if (s) g_override_shape = *s;
}

static bool shape_valid(const Lattice *lat, int prec, int nl, int nrhs, const HopShape &s) {
  const int slab = (int)(lat->Vh / lat->L[0]);
  if (s.nb != 1 && s.nb != 2 && s.nb != 4 && s.nb != 8 && s.nb != 16) return false;
  if (s.nb == 16 && (prec == P64 || nl != 16)) return false;
  if (s.nb > 1 && s.nb > nrhs) return false;
  if (s.block != 64 && s.block != 128 && s.block != 256 && !(s.block == 512 && s.stage)) return false;
  if (s.tile && (s.tile % s.block != 0 || slab % s.tile != 0 || s.tile >= slab)) return false;
  if (s.plain && (s.plain != 1 || prec == P64 || s.nb != 1 || s.stage || s.nt)) return false;
  if (s.stage) {
    if ((s.stage != 1 && s.stage != 2) || nl != 16 || s.nb > 8 || s.nt) return false;
    if (prec == P64 && s.nb != 2 && s.nb != 4) return false;
    const size_t el = prec == P16 ? 16 : prec == P64 ? 48 : 24;  // stored bytes per site and RHS
    if ((size_t)s.nb * s.block * el > 65536) return false;
    if (s.stage == 2 && (s.tile || slab % 64 || lat->L[0] % (s.block / 64))) return false;
  }
  return true;
}

static std::map<std::tuple<const Lattice *, int, int, int, int>, HopShape> g_tune_cache;

bool cached_hop_shape(Links &L, int prec, int nrhs, int red, HopShape *out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_override && shape_valid(L.lat, prec, L.nl, nrhs, g_override_shape)) {
    *out = g_override_shape;
    return true;
  }
  auto it = g_tune_cache.find(std::make_tuple((const Lattice *)L.lat, prec, L.nl, nrhs, red));
  if (it == g_tune_cache.end()) return false;
  *out = it->second;
  return true;
}

HopShape tune_hop(Links &L, int prec, int nrhs, const DevBatch &src, DevBatch &dst, int eo_out,
                  int red, const DevBatch *x, const DevBatch *xs) {
  auto &cache = g_tune_cache;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_override && shape_valid(L.lat, prec, L.nl, nrhs, g_override_shape)) return g_override_shape;
  auto key = std::make_tuple((const Lattice *)L.lat, prec, L.nl, nrhs, red);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  QOP_REQUIRE(red == RED_NONE || red == RED_OUT2 || (red == RED_CG && x && xs && prec != P64),
              "tune_hop: RED_CG needs x and xs");
  Lattice *lat = L.lat;
  LinkSet &ls = L.get(prec);
  hipStream_t st = stream();
  HopArgs a;
  a.g = lat->g;
  a.parity = eo_out;
  a.links = ls.parity_data(lat, eo_out);
  a.lscale = ls.parity_scale(lat, eo_out);
  a.src = src.slot(0);
  a.dst = dst.slot(0);
  a.sstride = a.dstride = a.xstride = 3 * lat->Vh;
  a.nstride = lat->Vh;
  a.n = nrhs;
  // the epilogue being tuned: scratch partials, and for the fused CG update
  // a zero alpha (dst and x then keep their values)
  double *scratch = nullptr;
  if (red != RED_NONE) {
    const size_t np = (size_t)nrhs * ((lat->Vh + 63) / 64) + nrhs;
    dev_alloc(&scratch, np * sizeof(double));
    QOP_CHECK_HIP(hipMemsetAsync(scratch, 0, np * sizeof(double), st));
    a.partial = scratch + nrhs;
  }
  if (red == RED_CG) {
    a.x = x->slot(0);
    a.xs = xs->slot(0);
    a.cg_alpha = scratch;
    a.cg_stride = 1;
  }
  const bool hasx = red == RED_CG;
  hipEvent_t e0, e1;
  QOP_CHECK_HIP(hipEventCreate(&e0));
  QOP_CHECK_HIP(hipEventCreate(&e1));
  HopShape best;
  float best_ms = 1e30f;
  const int nbs[] = {1, 2, 4, 8, 16};
  const int blocks[] = {64, 128, 256};
  const int tiles[] = {0, 512, 1024, 2048, 4096};
  const int slab = (int)(lat->Vh / lat->L[0]);
  const int reps = lat->Vh > 200000 ? 8 : 10;
  auto time_shape = [&](const HopShape &cand) {
    cand.apply(a);
    launch_hop_args(a, prec, ls.nl, cand.nb, cand.block, red, hasx, st);
    QOP_CHECK_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) launch_hop_args(a, prec, ls.nl, cand.nb, cand.block, red, hasx, st);
    QOP_CHECK_HIP(hipEventRecord(e1, st));
    QOP_CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    QOP_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; best = cand; }
  };
  // stage 1: RHS per thread and block size, with XCD remap, plain link loads,
  // RHS-innermost order and the plain site order
  for (int nb : nbs) {
    if (nb > 1 && nb > nrhs) continue;
    if (nb == 16 && (prec == P64 || L.nl != 16)) continue;
    for (int blk : blocks) time_shape(HopShape{nb, blk, 1, 0, nrhs > nb ? 1 : 0, 0});
  }
  // stage 2: the remaining switches for the best (RHS per thread, block)
  const int nb = best.nb, blk = best.block;
  for (int rm = 0; rm < 2; rm++)
    for (int nt = 0; nt < 2; nt++)
      for (int ri = 0; ri < 2; ri++)
        for (int tl : tiles) {
          if (ri && nrhs <= nb) continue;
          if (tl && (tl % blk != 0 || slab % tl != 0 || tl >= slab)) continue;
          time_shape(HopShape{nb, blk, rm, nt, ri, tl, 0});
        }
  // stage 3: LDS staging, for the best RHS per thread / order, with and
  // without the XCD remap: of the block's source rows at every block size and
  // x0 tile, and of its x0 planes (plane-blocked order) at 2, 4 and 8 planes
  // per block (plain link loads)
  if (L.nl == 16) {
    const HopShape b2 = best;
    for (int rm = 0; rm < 2; rm++) {
      for (int sblk : blocks)
        for (int tl : tiles) {
          HopShape c = b2;
          c.remap = rm;
          c.block = sblk;
          c.nt = 0;
          c.tile = tl;
          c.stage = 1;
          if (c.nb > 8) c.nb = 8;
          if (shape_valid(lat, prec, L.nl, nrhs, c)) time_shape(c);
        }
      for (int sblk : {128, 256, 512})
        for (int snb : {b2.nb, b2.nb / 2}) {
          HopShape c = b2;
          c.remap = rm;
          c.block = sblk;
          c.nt = 0;
          c.tile = 0;
          c.stage = 2;
          c.nb = snb > 8 ? 8 : snb < 1 ? 1 : snb;
          if (shape_valid(lat, prec, L.nl, nrhs, c)) time_shape(c);
        }
    }
  }
  QOP_CHECK_HIP(hipEventDestroy(e0));
  QOP_CHECK_HIP(hipEventDestroy(e1));
  if (scratch) {
    QOP_CHECK_HIP(hipStreamSynchronize(st));
    QOP_CHECK_HIP(hipFree(scratch));
  }
  if (options().verbose)
    printf("TUNE: prec %d nrhs %d -> nb %d block %d remap %d nt %d inner %d tile %d stage %d "
           "plain %d (%.3f ms/hop)\n", prec, nrhs, best.nb, best.block, best.remap, best.nt,
           best.rhs_inner, best.tile, best.stage, best.plain, best_ms / reps);
  cache[key] = best;
  return best;
}

// ---------------------------------------------------------------------------
BatchSolver::BatchSolver(Links &L, int eo, int capacity, int sloppy_prec, double delta,
                         hipStream_t st)
    : L_(L), lat_(L.lat), eo_(eo), cap_(capacity), sp_(sloppy_prec), delta_(delta),
      join_(options().ru_join),
      st_(st ? st : stream()) {
  try {
    QOP_REQUIRE(eo == 0 || eo == 1, "batch solver: subset must be even or odd");
    QOP_REQUIRE(capacity >= 1 && capacity <= 256, "batch solver: capacity 1..256");
    QOP_REQUIRE(sp_ == P32 || sp_ == P16, "batch solver: sloppy precision must be FP32 or FP16");
    nch_ = (int)((lat_->Vh + 63) / 64);
    devbatch_alloc(x64_, lat_, P64, cap_);
    devbatch_alloc(r64_, lat_, P64, cap_);
    devbatch_alloc(b64_, lat_, P64, cap_);
    devbatch_alloc(ax64_, lat_, P64, cap_);
    devbatch_alloc(t64_, lat_, P64, cap_);
    devbatch_alloc(rs_, lat_, sp_, cap_);
    devbatch_alloc(ps_, lat_, sp_, cap_);
    devbatch_alloc(ts_, lat_, sp_, cap_);
    devbatch_alloc(xs_, lat_, P32, cap_);
    dev_alloc(&rhs_, cap_ * sizeof(BsRhs));
    QOP_CHECK_HIP(hipHostMalloc(&rhs_h_, cap_ * sizeof(BsRhs), hipHostMallocDefault));
    dev_alloc(&active_, 2 * cap_ * sizeof(int));
    sel_ = active_ + cap_;
    QOP_CHECK_HIP(hipHostMalloc(&active_h_, 2 * cap_ * sizeof(int), hipHostMallocDefault));
    sel_h_ = active_h_ + cap_;
    dev_alloc(&ru_, (3 * cap_ + 1) * sizeof(int));
    ru_cont_ = ru_ + cap_;
    slots_ = ru_ + 2 * cap_;
    QOP_CHECK_HIP(hipMemsetAsync(ru_, 0, (3 * cap_ + 1) * sizeof(int), st_));
    dev_alloc(&glob_, sizeof(BsGlobal));
    QOP_CHECK_HIP(hipHostMalloc(&glob_h_, sizeof(BsGlobal), hipHostMallocDefault));
    dev_alloc(&partA_, 3 * (size_t)cap_ * nch_ * sizeof(double));
    partB_ = partA_ + (size_t)cap_ * nch_;
    partC_ = partB_ + (size_t)cap_ * nch_;
    dev_alloc(&coef_, (2 + 6) * (size_t)cap_ * sizeof(double));
    lcoef_ = coef_ + 2 * cap_;
    dev_alloc(&sums_, 2 * (size_t)cap_ * sizeof(double));
    QOP_CHECK_HIP(hipHostMalloc(&sums_h_, 2 * (size_t)cap_ * sizeof(double), hipHostMallocDefault));
    QOP_CHECK_HIP(hipHostMalloc(&hc_h_, 2 * (size_t)cap_ * sizeof(double), hipHostMallocDefault));
    memset(rhs_h_, 0, cap_ * sizeof(BsRhs));
    memset(active_h_, 0, 2 * cap_ * sizeof(int));
    glob_h_->status = 0;
    glob_h_->pending = 0;  // the status word stays 0: slots never pause the batch
    QOP_CHECK_HIP(hipMemcpyAsync(rhs_, rhs_h_, cap_ * sizeof(BsRhs), hipMemcpyHostToDevice, st_));
    QOP_CHECK_HIP(hipMemcpyAsync(active_, active_h_, 2 * cap_ * sizeof(int), hipMemcpyHostToDevice, st_));
    QOP_CHECK_HIP(hipMemcpyAsync(glob_, glob_h_, sizeof(BsGlobal), hipMemcpyHostToDevice, st_));
    loaded_.assign(cap_, 0);
    mass_.assign(cap_, 0.0);
    // Build the sloppy and FP64 link sets now: the iteration is captured into
    // a graph later, and a conversion (allocation, host sync) cannot run then.
    L_.get(sp_);
    L_.get(P64);
    // the two hops of the iteration are tuned with their own epilogues (the
    // second on rs_ / xs_ with alpha 0, while no slot is loaded)
    shape_ = tune_hop(L_, sp_, cap_, ps_, ts_, 1 - eo_, RED_OUT2);
    shape2_ = tune_hop(L_, sp_, cap_, ts_, rs_, eo_, RED_CG, &ps_, &xs_);
    ru_shape_ = tune_hop(L_, P64, cap_, x64_, t64_, 1 - eo_);

    L_.get(P64);
    // allocations are cleared and links converted on the library stream
    sync();
    QOP_CHECK_HIP(hipStreamSynchronize(st_));
  } catch (const DeviceOutOfMemory &) {
    release();  // a failed allocation frees what was built
    throw;
  }
}

BatchSolver::~BatchSolver() { release(); }

void BatchSolver::release() {
  QOP_CHECK_HIP(hipStreamSynchronize(st_));
  if (ev_) (void)hipEventDestroy(ev_);
  graph_.reset();
  for (DevBatch *b : {&x64_, &r64_, &b64_, &ax64_, &t64_, &rs_, &ps_, &ts_, &xs_}) devbatch_free(*b);
  (void)hipFree(rhs_);
  (void)hipHostFree(rhs_h_);
  (void)hipFree(active_);
  (void)hipFree(ru_);
  (void)hipHostFree(active_h_);
  (void)hipFree(glob_);
  (void)hipHostFree(glob_h_);
  (void)hipFree(partA_);
  (void)hipFree(coef_);
  (void)hipFree(sums_);
  (void)hipHostFree(sums_h_);
  (void)hipHostFree(hc_h_);
}

int BatchSolver::free_slot() const {
  for (int i = 0; i < cap_; i++)
    if (!loaded_[i]) return i;
  return -1;
}

int BatchSolver::nactive() const {
  int n = 0;
  for (int i = 0; i < cap_; i++) n += active_h_[i];
  return n;
}

// helper: lincomb over the batch with per-slot selection
static void batch_lincomb(Lattice *lat, int cap, const DevBatch &out, int nin, const DevBatch *const *in,
                          const double *cr, const double *ci, const double *dev_coef, const int *sel,
                          double *partial, hipStream_t st) {
  LincombArgs a;
  a.Vh = (int)lat->Vh;
  a.n = cap;
  a.out = out.slot(0);
  a.ostride = 3 * lat->Vh;
  a.nstride = lat->Vh;
  int pin[3];
  for (int i = 0; i < nin; i++) {
    a.v[i] = in[i]->slot(0);
    a.stride[i] = 3 * lat->Vh;
    if (cr) { a.cr[i] = cr[i]; a.ci[i] = ci ? ci[i] : 0.0; }
    pin[i] = in[i]->prec;
  }
  a.coef = dev_coef;
  a.active = sel;
  a.partial = partial;
  launch_lincomb(a, out.prec, pin, nin, partial != nullptr, st);
}

void BatchSolver::apply_A64(const int *mask, bool listed) {
  hipStream_t st = st_;
  LinkSet &ls = L_.get(P64);
  const int opp = 1 - eo_;
  HopArgs a;
  a.g = lat_->g;
  a.parity = opp;
  a.links = ls.parity_data(lat_, opp);
  a.lscale = ls.parity_scale(lat_, opp);
  a.src = x64_.slot(0);
  a.dst = t64_.slot(0);
  a.sstride = a.dstride = a.xstride = 3 * lat_->Vh;
  a.nstride = lat_->Vh;
  a.n = cap_;
  // the flagged slots, compacted into consecutive RHS of the launch, so k of
  // them take ceil(k / RHS per thread) link reads wherever they sit in the
  // batch; the groups past them read nothing
  if (!listed) bs_launch_compact(mask, cap_, slots_, st);
  a.slots = slots_;
  ru_shape_.apply(a);
  launch_hop_args(a, P64, ls.nl, ru_shape_.nb, ru_shape_.block, RED_NONE, false, st);
  HopArgs b = a;
  b.parity = eo_;
  b.links = ls.parity_data(lat_, eo_);
  b.lscale = ls.parity_scale(lat_, eo_);
  b.src = t64_.slot(0);
  b.dst = ax64_.slot(0);
  b.x = x64_.slot(0);
  b.coef = coef_;
  launch_hop_args(b, P64, ls.nl, ru_shape_.nb, ru_shape_.block, RED_NONE, true, st);
}

void BatchSolver::load(int slot, const DevVec &src, const DevVec *x0, double mass, double tol,
                       int max_iter) {
  const LoadReq q{slot, &src, x0, mass, tol, max_iter};
  load_many(&q, 1);
}

// Several slots in one pass: the FP64 operator for the initial residuals
// b - A x0 is applied once, masked to the slots that have an initial guess
// (each slot's arithmetic is its own, so the results do not depend on which
// slots load together).
void BatchSolver::load_many(const LoadReq *req, int nreq) {
  if (nreq <= 0) return;
  hipStream_t st = st_;
  const long Vh = lat_->Vh;
  auto copy_half = [&](VecView dst, int pdst, VecView s, int ps_) {
    LincombArgs a;
    a.Vh = (int)Vh;
    a.n = 1;
    a.out = dst;
    a.v[0] = s;
    a.cr[0] = 1.0;
    int pin[1] = {ps_};
    launch_lincomb(a, pdst, pin, 1, false, st);
  };
  for (int i = 0; i < cap_; i++) sel_h_[i] = 0;
  bool any_x0 = false;
  for (int k = 0; k < nreq; k++) {
    const LoadReq &q = req[k];
    QOP_REQUIRE(q.slot >= 0 && q.slot < cap_ && !loaded_[q.slot] && !sel_h_[q.slot],
                "batch solver: slot not free");
    QOP_REQUIRE(q.src->lat == lat_, "batch solver: lattice mismatch");
    copy_half(b64_.slot(q.slot), P64, q.src->half(eo_), P64);
    if (q.x0) copy_half(x64_.slot(q.slot), P64, q.x0->half(eo_), P64);
    else QOP_CHECK_HIP(hipMemsetAsync(x64_.slot(q.slot).data, 0, 3 * Vh * sizeof(double2), st));
    // hop coefficients of this slot: A = m^2 - D D
    hc_h_[2 * k] = -1.0;
    hc_h_[2 * k + 1] = q.mass * q.mass;
    QOP_CHECK_HIP(hipMemcpyAsync(coef_ + 2 * q.slot, hc_h_ + 2 * k, 2 * sizeof(double),
                                 hipMemcpyHostToDevice, st));
    sel_h_[q.slot] = 1;
    any_x0 |= q.x0 != nullptr;
  }
  QOP_CHECK_HIP(hipMemcpyAsync(sel_, sel_h_, cap_ * sizeof(int), hipMemcpyHostToDevice, st));
  // r = b - A x (x = 0 for the slots without a guess: A x = 0 there)
  if (any_x0) {
    apply_A64(sel_);
    const DevBatch *in[2] = {&b64_, &ax64_};
    double cr[2] = {1.0, -1.0};
    batch_lincomb(lat_, cap_, r64_, 2, in, cr, nullptr, nullptr, sel_, partC_, st);
  } else {
    const DevBatch *in[1] = {&b64_};
    double cr[1] = {1.0};
    batch_lincomb(lat_, cap_, r64_, 1, in, cr, nullptr, nullptr, sel_, partC_, st);
  }
  bs_launch_sum(partC_, nch_, cap_, sel_, sums_, st);
  // |b|^2
  {
    DotArgs d;
    d.Vh = (int)Vh;
    d.n = cap_;
    d.u = b64_.slot(0);
    d.ustride = 3 * Vh;
    d.nstride = Vh;
    d.partial = partA_;
    d.active = sel_;
    launch_dot(d, P64, P64, 0, st);
    bs_launch_sum(partA_, nch_, cap_, sel_, sums_ + cap_, st);
  }
  // rs = r, ps = rs, xs = 0
  {
    const DevBatch *in[1] = {&r64_};
    double cr[1] = {1.0};
    batch_lincomb(lat_, cap_, rs_, 1, in, cr, nullptr, nullptr, sel_, nullptr, st);
    const DevBatch *in2[1] = {&rs_};
    batch_lincomb(lat_, cap_, ps_, 1, in2, cr, nullptr, nullptr, sel_, nullptr, st);
    // |ps|^2 partials of the loaded slots, consumed by the first iteration's p.Ap
    DotArgs d;
    d.Vh = (int)Vh;
    d.n = cap_;
    d.u = ps_.slot(0);
    d.ustride = 3 * Vh;
    d.nstride = Vh;
    d.partial = partB_;
    d.active = sel_;
    launch_dot(d, sp_, sp_, 0, st);
    for (int k = 0; k < nreq; k++)
      QOP_CHECK_HIP(hipMemsetAsync(xs_.slot(req[k].slot).data, 0, 3 * Vh * sizeof(float2), st));
  }
  QOP_CHECK_HIP(hipMemcpyAsync(sums_h_, sums_, 2 * cap_ * sizeof(double), hipMemcpyDeviceToHost, st));
  QOP_CHECK_HIP(hipStreamSynchronize(st_));
  for (int k = 0; k < nreq; k++) {
    const LoadReq &q = req[k];
    const int slot = q.slot;
    BsRhs &R = rhs_h_[slot];
    memset(&R, 0, sizeof R);
    R.m2 = q.mass * q.mass;
    R.r2 = sums_h_[slot];
    R.r2_rel = R.r2;
    R.b2 = sums_h_[cap_ + slot];
    R.stop2 = q.tol * q.tol * R.b2;
    R.max_iter = q.max_iter;
    R.delta2 = delta_ * delta_;
    R.join2 = join_ * join_;
    mass_[slot] = q.mass;
    loaded_[slot] = 1;
    const bool done = R.r2 <= R.stop2;
    if (done) R.reason = 1;
    active_h_[slot] = done ? 0 : 1;
    QOP_CHECK_HIP(hipMemcpyAsync(rhs_ + slot, &R, sizeof R, hipMemcpyHostToDevice, st));
    QOP_CHECK_HIP(hipMemcpyAsync(active_ + slot, active_h_ + slot, sizeof(int), hipMemcpyHostToDevice, st));
  }
  QOP_CHECK_HIP(hipStreamSynchronize(st_));
}

void BatchSolver::enqueue_iteration() {
  hipStream_t st = st_;
  LinkSet &ls = L_.get(sp_);
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
  a.status = &glob_->status;
  a.partial = partA_;
  shape_.apply(a);
  launch_hop_args(a, sp_, ls.nl, shape_.nb, shape_.block, RED_OUT2, false, st);
  // second hop with the CG update fused: r -= alpha (m^2 p - D t), x += alpha p
  HopArgs b = a;
  b.parity = eo_;
  b.links = ls.parity_data(lat_, eo_);
  b.lscale = ls.parity_scale(lat_, eo_);
  b.src = ts_.slot(0);
  b.dst = rs_.slot(0);
  b.x = ps_.slot(0);
  b.xs = xs_.slot(0);
  b.coef = coef_;
  b.cg_alpha = (const double *)((const char *)rhs_ + offsetof(BsRhs, alpha));
  b.cg_stride = (int)(sizeof(BsRhs) / sizeof(double));
  b.partial = partC_;
  bs_launch_scalar1(rhs_, active_, &glob_->status, partA_, partB_, nch_, cap_, ru_cont_, st);
  shape2_.apply(b);
  launch_hop_args(b, sp_, ls.nl, shape2_.nb, shape2_.block, RED_CG, true, st);
  bs_launch_scalar2(rhs_, active_, glob_, partC_, nch_, cap_, ru_, st);
  if (join_ > 0) bs_launch_ru_join(rhs_, active_, ru_, cap_, st);
  bs_launch_pupdate(rhs_, active_, &glob_->status, sp_, (int)Vh, rs_.slot(0), ps_.slot(0), 3 * Vh, cap_, 1,
                    partB_, st);
  // reliable update of the flagged slots, all on the device, over their
  // compacted list
  const long stride = 3 * Vh;
  bs_launch_compact(ru_, cap_, slots_, st);
  bs_launch_ru(rhs_, active_, ru_, ru_cont_, slots_, lcoef_, sp_, (int)Vh, x64_.slot(0), xs_.slot(0), b64_.slot(0),
               ax64_.slot(0), rs_.slot(0), ps_.slot(0), stride, partA_, partB_, partC_, nch_, cap_, 0, partB_, st);
  apply_A64(ru_, true);
  bs_launch_ru(rhs_, active_, ru_, ru_cont_, slots_, lcoef_, sp_, (int)Vh, x64_.slot(0), xs_.slot(0), b64_.slot(0),
               ax64_.slot(0), rs_.slot(0), ps_.slot(0), stride, partA_, partB_, partC_, nch_, cap_, 1, partB_, st);
  bs_launch_ru(rhs_, active_, ru_, ru_cont_, slots_, lcoef_, sp_, (int)Vh, x64_.slot(0), xs_.slot(0), b64_.slot(0),
               ax64_.slot(0), rs_.slot(0), ps_.slot(0), stride, partA_, partB_, partC_, nch_, cap_, 2, partB_, st);
  bs_launch_ru(rhs_, active_, ru_, ru_cont_, slots_, lcoef_, sp_, (int)Vh, x64_.slot(0), xs_.slot(0), b64_.slot(0),
               ax64_.slot(0), rs_.slot(0), ps_.slot(0), stride, partA_, partB_, partC_, nch_, cap_, 3, partB_, st);
}

void BatchSolver::build_graph() {
  const int chunk = options().chunk > 0 ? options().chunk : 16;
  if (graph_ && graph_chunk_ == chunk) return;
  graph_.begin(st_);
  for (int c = 0; c < chunk; c++) enqueue_iteration();
  graph_.end(st_);
  graph_chunk_ = chunk;
}

void BatchSolver::launch_chunk() {
  QOP_REQUIRE(!inflight_, "batch solver: a chunk is already in flight");
  bool any = false;
  for (int i = 0; i < cap_; i++) any |= active_h_[i] != 0;
  if (!any) return;
  if (options().use_graph) {
    build_graph();
    graph_.launch(st_);
  } else {
    const int chunk = options().chunk > 0 ? options().chunk : 16;
    for (int c = 0; c < chunk; c++) enqueue_iteration();
  }
  // slots finish on the device (bs_ru_scalar clears active[b])
  QOP_CHECK_HIP(hipMemcpyAsync(active_h_, active_, cap_ * sizeof(int), hipMemcpyDeviceToHost, st_));
  if (!ev_) QOP_CHECK_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  QOP_CHECK_HIP(hipEventRecord(ev_, st_));
  inflight_ = true;
}

bool BatchSolver::chunk_done() {
  if (!inflight_) return true;
  hipError_t e = hipEventQuery(ev_);
  if (e == hipErrorNotReady) return false;
  QOP_CHECK_HIP(e);
  return true;
}

void BatchSolver::collect(std::vector<int> &done) {
  if (inflight_) {
    QOP_CHECK_HIP(hipEventSynchronize(ev_));
    inflight_ = false;
  }
  for (int i = 0; i < cap_; i++)
    if (loaded_[i] == 1 && !active_h_[i]) {
      loaded_[i] = 2;  // finished, not yet released
      done.push_back(i);
    }
}

void BatchSolver::step(std::vector<int> &done) {
  launch_chunk();
  collect(done);
}

void BatchSolver::run_all(std::vector<int> &done) {
  while (nactive() > 0) step(done);
  step(done);
}

void BatchSolver::result(int slot, DevVec &out, bool times_mass, SlotResult *res) {
  QOP_REQUIRE(slot >= 0 && slot < cap_ && loaded_[slot], "batch solver: slot not loaded");
  hipStream_t st = st_;
  LincombArgs a;
  a.Vh = (int)lat_->Vh;
  a.n = 1;
  a.out = out.half(eo_);
  a.v[0] = x64_.slot(slot);
  a.cr[0] = times_mass ? mass_[slot] : 1.0;
  int pin[1] = {P64};
  launch_lincomb(a, out.prec, pin, 1, false, st);
  QOP_CHECK_HIP(hipStreamSynchronize(st_));
  if (res) {
    QOP_CHECK_HIP(hipMemcpy(rhs_h_ + slot, rhs_ + slot, sizeof(BsRhs), hipMemcpyDeviceToHost));
    const BsRhs &R = rhs_h_[slot];
    res->iter = R.iter;
    res->nupdate = R.nupdate;
    res->reason = R.reason;
    res->true_rsq = R.b2 > 0 ? R.r2_rel / R.b2 : 0.0;
  }
}

void BatchSolver::release(int slot) {
  QOP_REQUIRE(slot >= 0 && slot < cap_, "batch solver: bad slot");
  loaded_[slot] = 0;
  active_h_[slot] = 0;
  QOP_CHECK_HIP(hipMemcpyAsync(active_ + slot, active_h_ + slot, sizeof(int), hipMemcpyHostToDevice, st_));
}

// ---------------------------------------------------------------------------
// Solver dispatch for the QOP entry points: the reference algorithm by
// default; with option qmx_solver = 1, EVEN/ODD single-mass solves use the
// reliable-update mixed-precision solver and multi-mass solves use one batch
// of per-mass systems.
void invert_dispatch(Links &L, InvArg *ia, ResArg *ra, double mass, DevVec &out, const DevVec &in,
                     double *flop, double *sec) {
  const Options &o = options();
  if (o.solver == 0 || ia->evenodd == 2 || o.inner_prec == P64) {
    invert(L, ia, ra, mass, out, in, flop, sec);
    return;
  }
  double t0 = wall_time();
  BatchSolver bs(L, ia->evenodd, 1, o.inner_prec, o.delta);
  double tol = std::sqrt(ra->rsqmin);
  bs.load(0, in, nullptr, mass, tol, ia->max_iter);
  std::vector<int> done;
  bs.run_all(done);
  SlotResult r;
  bs.result(0, out, true, &r);
  ra->final_iter = r.iter;
  ra->final_rsq = r.true_rsq;
  ra->final_rel = 0;
  ra->final_restart = r.nupdate;
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl);
  if (flop) *flop = nflop * r.iter * (double)L.lat->V;
  if (sec) *sec = wall_time() - t0;
}

void invert_multi_dispatch(Links &L, InvArg *ia, ResArg **ra, const double *masses, const int *nmass,
                           DevVec **out, DevVec *const *in, int nsrc, double *flop, double *sec) {
  const Options &o = options();
  if (o.solver == 0 || ia->evenodd == 2 || o.inner_prec == P64) {
    invert_multi(L, ia, ra, masses, nmass, out, in, nsrc, flop, sec);
    return;
  }
  double t0 = wall_time();
  int nsolve = 0;
  for (int i = 0; i < nsrc; i++) nsolve += nmass[i];
  const int cap = std::min(nsolve, 16);
  BatchSolver bs(L, ia->evenodd, cap, o.inner_prec, o.delta);
  std::vector<int> slot_of(cap, -1);
  std::vector<int> src_of(nsolve);
  {
    int k = 0;
    for (int i = 0; i < nsrc; i++)
      for (int j = 0; j < nmass[i]; j++) src_of[k++] = i;
  }
  int next = 0, finished = 0;
  long iters = 0;
  std::vector<BatchSolver::LoadReq> reqs;
  while (finished < nsolve) {
    reqs.clear();
    for (int s = 0; s < bs.capacity() && next < nsolve; s++) {
      if (!bs.slot_free(s)) continue;
      reqs.push_back({s, in[src_of[next]], nullptr, masses[next], std::sqrt(ra[next]->rsqmin),
                      ia->max_iter});
      slot_of[s] = next++;
    }
    bs.load_many(reqs.data(), (int)reqs.size());
    std::vector<int> done;
    bs.step(done);
    for (int d : done) {
      int k = slot_of[d];
      SlotResult r;
      bs.result(d, *out[k], true, &r);
      ra[k]->final_iter = r.iter;
      ra[k]->final_rsq = r.true_rsq;
      ra[k]->final_rel = 0;
      ra[k]->final_restart = r.nupdate;
      iters += r.iter;
      bs.release(d);
      finished++;
    }
  }
  double nflop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl);
  if (flop) *flop = nflop * (double)iters * (double)L.lat->V;
  if (sec) *sec = wall_time() - t0;
}

}  // namespace qop
