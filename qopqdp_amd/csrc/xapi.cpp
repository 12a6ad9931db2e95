// Extension entry points (QOPX_*): batched mixed-precision solves, hop
// timing for roofline measurements, and link-precision control.  None of
// these has a counterpart in the reference; they sit beside the QOP ABI.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "batch.h"
#include "handles.h"

using namespace qop;

// the body of QOPX_D3_asqtad_hop_time (below)
static double hop_time_impl(QOP_D3_FermionLinksAsqtad *fla, int prec, int nrhs, int nrep,
                            int red, int *shape) {
  Links &L = fla->L;
  Lattice *lat = L.lat;
  int p = prec == 64 ? P64 : prec == 32 ? P32 : P16;
  QOP_REQUIRE(nrhs >= 1 && nrhs <= 64 && nrep >= 1, "hop_time: bad arguments");
  DevBatch src, dst, xb, xs;
  DevVec tmp;
  // everything is freed on return, also when an allocation fails
  struct Free {
    DevBatch *b[4];
    DevVec *v;
    ~Free() {
      for (DevBatch *x : b) devbatch_free(*x);
      devvec_free(*v);
    }
  } fr{{&src, &dst, &xb, &xs}, &tmp};
  devbatch_alloc(src, lat, p, nrhs);
  devbatch_alloc(dst, lat, p, nrhs);
  // fill the source with a constant so FP16 norms are nonzero
  {
    std::vector<double> one((size_t)lat->V * 6, 0.5);
    devvec_alloc(tmp, lat, P64);
    DevBuf<double> dbuf(one.size() * sizeof(double));
    double *d = dbuf.p;
    QOP_CHECK_HIP(hipMemcpy(d, one.data(), one.size() * sizeof(double), hipMemcpyHostToDevice));
    launch_raw_to_vec(lat, d, tmp, 2, stream());
    for (int b = 0; b < nrhs; b++) {
      LincombArgs a;
      a.Vh = (int)lat->Vh;
      a.out = src.slot(b);
      a.v[0] = tmp.half(0);
      a.cr[0] = 1.0;
      int pin[1] = {P64};
      launch_lincomb(a, p, pin, 1, false, stream());
    }
    sync();
  }
  QOP_REQUIRE(red >= 0 && red <= 2 && (red != 2 || p != P64), "hop_time: bad reduction");
  if (red == 2) {  // operands of the fused CG update: p (x), x_FP32
    devbatch_alloc(xb, lat, p, nrhs);
    devbatch_alloc(xs, lat, P32, nrhs);
    sync();
  }
  // the shape the batched solver tunes for this epilogue
  const int rk = red == 2 ? RED_CG : red == 1 ? RED_OUT2 : RED_NONE;
  HopShape hs = tune_hop(L, p, nrhs, src, dst, 1, rk, red == 2 ? &xb : nullptr,
                         red == 2 ? &xs : nullptr);
  LinkSet &ls = L.get(p);
  HopArgs a;
  a.g = lat->g;
  a.parity = 1;
  a.links = ls.parity_data(lat, 1);
  a.lscale = ls.parity_scale(lat, 1);
  a.src = src.slot(0);
  a.dst = dst.slot(0);
  a.sstride = a.dstride = 3 * lat->Vh;
  a.nstride = lat->Vh;
  a.n = nrhs;
  hs.apply(a);
  DevBuf<double> part, alpha_buf;
  if (red) {
    dev_alloc(&part.p, (size_t)nrhs * ((lat->Vh + 63) / 64) * sizeof(double));
    a.partial = part.p;
  }
  double *alpha = nullptr;
  if (red == 2) {  // the fused CG update: x = p, dst = r, alpha 0
    dev_alloc(&alpha_buf.p, nrhs * sizeof(double));
    alpha = alpha_buf.p;
    QOP_CHECK_HIP(hipMemsetAsync(alpha, 0, nrhs * sizeof(double), stream()));
    QOP_CHECK_HIP(hipMemcpyAsync(xb.data, src.data, (size_t)nrhs * 3 * lat->Vh * prec_cbytes(p),
                                 hipMemcpyDeviceToDevice, stream()));
    if (p == P16)
      QOP_CHECK_HIP(hipMemcpyAsync(xb.norm, src.norm, (size_t)nrhs * lat->Vh * sizeof(float),
                                   hipMemcpyDeviceToDevice, stream()));
    a.x = xb.slot(0);
    a.xstride = 3 * lat->Vh;
    a.xs = xs.slot(0);
    a.cg_alpha = alpha;
    a.cg_stride = 1;
    a.alpha = -1.0;
    a.beta = 1e-4;
    sync();
  }
  hipStream_t st = stream();
  for (int w = 0; w < 3; w++) launch_hop_args(a, p, ls.nl, hs.nb, hs.block, rk, red == 2, st);
  hipEvent_t e0, e1;
  QOP_CHECK_HIP(hipEventCreate(&e0));
  QOP_CHECK_HIP(hipEventCreate(&e1));
  QOP_CHECK_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < nrep; r++) launch_hop_args(a, p, ls.nl, hs.nb, hs.block, rk, red == 2, st);
  QOP_CHECK_HIP(hipEventRecord(e1, st));
  QOP_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QOP_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  QOP_CHECK_HIP(hipEventDestroy(e0));
  QOP_CHECK_HIP(hipEventDestroy(e1));
  sync();
  if (shape) {
    shape[0] = hs.nb; shape[1] = hs.block; shape[2] = hs.remap; shape[3] = hs.nt; shape[4] = hs.rhs_inner;
    shape[5] = hs.tile;
    shape[6] = hs.stage;
    shape[7] = hs.plain;
  }
  return ms / nrep;
}

// the body of QOPX_D3_asqtad_hop (below)
static void hop_impl(QOP_D3_FermionLinksAsqtad *fla, int prec, int n, QOP_D3_ColorVector *out[],
                     QOP_D3_ColorVector *in[], int eo_out) {
  Links &L = fla->L;
  Lattice *lat = L.lat;
  QOP_REQUIRE(prec == 64 || prec == 32 || prec == 16, "hop: precision must be 64, 32 or 16");
  QOP_REQUIRE(n >= 1 && n <= 64, "hop: 1..64 vectors");
  QOP_REQUIRE(eo_out == QOP_EVEN || eo_out == QOP_ODD, "hop: output subset must be EVEN or ODD");
  const int p = prec == 64 ? P64 : prec == 32 ? P32 : P16;
  const int src_par = 1 - eo_out;
  DevBatch src, dst;
  struct Free {
    DevBatch *b[2];
    ~Free() {
      for (DevBatch *x : b) devbatch_free(*x);
    }
  } fr{{&src, &dst}};
  devbatch_alloc(src, lat, p, n);
  devbatch_alloc(dst, lat, p, n);
  int pin[1] = {P64};
  for (int b = 0; b < n; b++) {  // FP64 -> stored precision (FP16: per-site norm)
    LincombArgs a;
    a.Vh = (int)lat->Vh;
    a.out = src.slot(b);
    a.v[0] = in[b]->v.half(src_par);
    a.cr[0] = 1.0;
    launch_lincomb(a, p, pin, 1, false, stream());
  }
  HopShape hs = tune_hop(L, p, n, src, dst, eo_out, RED_NONE);
  LinkSet &ls = L.get(p);
  HopArgs a;
  a.g = lat->g;
  a.parity = eo_out;
  a.links = ls.parity_data(lat, eo_out);
  a.lscale = ls.parity_scale(lat, eo_out);
  a.src = src.slot(0);
  a.dst = dst.slot(0);
  a.sstride = a.dstride = 3 * lat->Vh;
  a.nstride = lat->Vh;
  a.n = n;
  hs.apply(a);
  launch_hop_args(a, p, ls.nl, hs.nb, hs.block, RED_NONE, false, stream());
  int pd[1] = {p};
  for (int b = 0; b < n; b++) {  // stored precision -> FP64
    LincombArgs c;
    c.Vh = (int)lat->Vh;
    c.out = out[b]->v.half(eo_out);
    c.v[0] = dst.slot(b);
    c.cr[0] = 1.0;
    launch_lincomb(c, P64, pd, 1, false, stream());
  }
  sync();
}

extern "C" {

// The hop out = D_{eo_out, 1-eo_out} in of the sloppy solvers, in precision
// prec (64/32/16: FP16 vectors with a per-site norm and FP16 links with a
// per-link scale) on n right-hand sides in one multi-RHS launch with the
// autotuned shape (the kernel the batched solvers run): each in[i] is
// rounded to the stored precision on the source parity, the result is
// written to out[i] on eo_out in FP64.  Returns 0, or -1 when the device
// memory is exhausted.
int QOPX_D3_asqtad_hop(QOP_D3_FermionLinksAsqtad *fla, int prec, int n, QOP_D3_ColorVector *out[],
                       QOP_D3_ColorVector *in[], QOP_evenodd_t eo_out) {
  return guard_oom(__func__, nullptr, [&] { hop_impl(fla, prec, n, out, in, eo_out); }) ? 0 : -1;
}

// Time `nrep` launches of the even->odd hop kernel in precision prec
// (64/32/16) on nrhs right-hand sides with the autotuned launch shape.
// red = 1 selects the instance with the fused |out|^2 reduction (the first
// hop of a batched CG iteration), red = 2 the instance with the fused CG
// update r -= a (m^2 p - D t), x_FP32 += a p, |r|^2 (the second hop; sloppy
// precisions only).  Returns the mean ms per launch from
// HIP events on the library stream; shape[8] receives (rhs per thread, block,
// XCD remap, non-temporal links, RHS-innermost block order, x0 tile, LDS
// staging, plain loads).
double QOPX_D3_asqtad_hop_time(QOP_D3_FermionLinksAsqtad *fla, int prec, int nrhs, int nrep,
                               int red, int *shape) {
  double ms = -1.0;  // -1: the device memory is exhausted
  guard_oom(__func__, nullptr, [&] { ms = hop_time_impl(fla, prec, nrhs, nrep, red, shape); });
  return ms;
}

// Force the multi-RHS hop launch shape (rhs per thread, block, XCD remap,
// non-temporal links, RHS-innermost order, x0 tile, LDS staging, plain loads)
// for solvers built from now on, wherever it is valid; null restores
// autotuning.
void QOPX_hop_shape_override(const int *shape) {
  if (!shape) {
    set_hop_shape_override(nullptr);
    return;
  }
  HopShape s{shape[0], shape[1], shape[2], shape[3], shape[4], shape[5], shape[6], shape[7]};
  set_hop_shape_override(&s);
}

// Batched mixed-precision solves of n EVEN (or ODD) systems
// out[i] = m_i A_i^-1 in[i] (the QOP m/A normalisation), each to its own
// FP64 true relative residual sqrt(res[i]->rsqmin).  Up to `capacity`
// systems iterate together; a finished slot is refilled with the next system
// (continuous batching).  prec: sloppy precision 32 or 16.  latency[i]
// (optional) receives the seconds from the call's start to system i's
// completion.  Returns the number of systems that reached their tolerance.
int QOPX_D3_asqtad_invert_batch(QOP_info_t *info, QOP_D3_FermionLinksAsqtad *fla,
                                QOP_invert_arg_t *inv_arg, QOP_resid_arg_t *res[], double masses[],
                                QOP_D3_ColorVector *out[], QOP_D3_ColorVector *in[], int n,
                                int capacity, int prec, double *latency) {
  Links &L = fla->L;
  QOP_REQUIRE(inv_arg->evenodd == QOP_EVEN || inv_arg->evenodd == QOP_ODD,
              "invert_batch: subset must be EVEN or ODD");
  QOP_REQUIRE(n >= 0 && capacity >= 1, "invert_batch: bad sizes");
  if (n == 0) return 0;
  QOP_REQUIRE(prec == 16 || prec == 32, "invert_batch: sloppy precision must be 16 or 32");
  const int sp = prec == 16 ? P16 : P32;
  double t0 = wall_time();
  const int cap = std::min(capacity, n);
  int ok = 0;
  long iters = 0;
  const bool fit = guard_oom(__func__, info, [&] {
    BatchSolver bs(L, inv_arg->evenodd, cap, sp, options().delta);
    std::vector<int> slot_of(cap, -1);
    int next = 0, finished = 0;
    std::vector<BatchSolver::LoadReq> reqs;
    while (finished < n) {
      reqs.clear();
      for (int s = 0; s < cap && next < n; s++) {
        if (!bs.slot_free(s)) continue;
        reqs.push_back({s, &in[next]->v, nullptr, masses[next], std::sqrt(res[next]->rsqmin),
                        inv_arg->max_iter});
        slot_of[s] = next++;
      }
      bs.load_many(reqs.data(), (int)reqs.size());
      std::vector<int> done;
      bs.step(done);
      for (int d : done) {
        int k = slot_of[d];
        SlotResult r;
        bs.result(d, out[k]->v, true, &r);
        res[k]->final_iter = r.iter;
        res[k]->final_rsq = r.true_rsq;
        res[k]->final_rel = 0;
        res[k]->final_restart = r.nupdate;
        ok += r.reason == 1;
        iters += r.iter;
        if (latency) latency[k] = wall_time() - t0;
        bs.release(d);
        finished++;
      }
    }
  });
  if (!fit) return -1;
  if (info) {
    info->final_sec = wall_time() - t0;
    info->final_flop = 0.5 * (20 * 3 + 2 * 8 * 9 * L.nl) * (double)iters * (double)L.lat->V;
    info->status = QOP_SUCCESS;
  }
  return ok;
}

// Mean ms per launch of a plain device-to-device copy of `bytes` (read +
// write), used to calibrate HBM traffic counters and report the measured
// copy bandwidth next to the nominal peak.
double QOPX_copy_time(unsigned long long bytes, int nrep) {
  QOP_REQUIRE(bytes >= 16 && bytes % 16 == 0 && nrep >= 1, "copy_time: bytes must be a multiple of 16");
  DevBuf<void> abuf, bbuf;
  if (!guard_oom(__func__, nullptr, [&] {
        dev_alloc(&abuf.p, bytes);
        dev_alloc(&bbuf.p, bytes);
      }))
    return -1.0;
  void *a = abuf.p, *b = bbuf.p;
  QOP_CHECK_HIP(hipMemsetAsync(a, 0, bytes, stream()));
  hipStream_t st = stream();
  for (int w = 0; w < 2; w++) launch_copy(a, b, bytes, st);
  hipEvent_t e0, e1;
  QOP_CHECK_HIP(hipEventCreate(&e0));
  QOP_CHECK_HIP(hipEventCreate(&e1));
  QOP_CHECK_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < nrep; r++) launch_copy(a, b, bytes, st);
  QOP_CHECK_HIP(hipEventRecord(e1, st));
  QOP_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QOP_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  QOP_CHECK_HIP(hipEventDestroy(e0));
  QOP_CHECK_HIP(hipEventDestroy(e1));
  return ms / nrep;
}

}  // extern "C"
