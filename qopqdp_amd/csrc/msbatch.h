// Batched multi-shift CG over several sources sharing one set of links.
//
// For every source b (a slot) and shift i the system (A_0 + s_i) x_bi = src_b,
// A_0 = m_0^2 - D_eo D_oe, s_i = m_i^2 - m_0^2 >= 0, is solved in the Krylov
// space of the base system with the zeta/beta/alpha recurrences of the
// reference multi-shift CG (lib/invert_cg_p.c:130-402), kept in FP64 per
// source.  Vectors are in a sloppy precision (FP32, or FP16 with per-site
// norms) with FP32 solution accumulators and FP64 reductions; a shift stops
// being updated once its estimated residual zeta_i^2 |r|^2 reaches its
// target.  Reliable updates: whenever a source's iterated residual has
// dropped by `delta` since its last update, every shift's FP32 accumulator is
// folded into its FP64 solution and the base residual is recomputed in FP64,
// r = b - A_0 x_0, and replaces the sloppy one (the shifted residuals stay
// zeta_i r).  The multi-shift solutions then seed FP64 reliable-update
// refinements (BatchSolver), so every (source, mass) reaches its FP64 true
// residual.
#pragma once
#include "batch.h"

namespace qop {

constexpr int MS_MAXSHIFT = 16;

struct MsSlot {
  int ns, imin, iter, max_iter;
  int done;          // finished (set by the iteration after the last update)
  int done_pending;  // every shift reached its target in this iteration
  int live[MS_MAXSHIFT];       // shift updated in this iteration
  int live_next[MS_MAXSHIFT];  // shift updated in the next iteration
  double r2, r2_old, pAp, b2;
  double m2;         // base mass^2
  double beta0;      // b[imin] of the current iteration (the hop epilogue reads it)
  double r2_max;     // |r|^2 at the last reliable update
  double delta2;     // reliable-update threshold (delta^2)
  int nupdate, pad_;
  double stop2[MS_MAXSHIFT];  // per-shift sloppy target on zeta^2 |r|^2
  double shift[MS_MAXSHIFT];
  double a[MS_MAXSHIFT], b[MS_MAXSHIFT], bo[MS_MAXSHIFT];
  double z[MS_MAXSHIFT], zo[MS_MAXSHIFT], zn[MS_MAXSHIFT];
};

void ms_launch_scalarA(MsSlot *s, const int *active, const double *partA, const double *partB,
                       int nch, int n, hipStream_t st);
void ms_launch_scalarB(MsSlot *s, int *active, const double *partC, int nch, int n, int *ru,
                       hipStream_t st);
// Reliable update of the slots flagged in ru: fold x64[k] += xs[k], xs[k] = 0
// for every shift (k = shift * cap + slot); residual r = b - ax into rs with
// |r|^2 partials; scalars (r2, r2_max, clear the flag).
void ms_launch_ru_fold(const int *ru, int Vh, VecView x64, VecView xs, long stride, int cap, int ns,
                       hipStream_t st);
void ms_launch_ru_resid(const int *ru, int prec, int Vh, VecView b64, VecView ax64, VecView rs,
                        long stride, int cap, double *part, hipStream_t st);
void ms_launch_ru_scalar(MsSlot *s, int *ru, const double *part, int nch, int n, hipStream_t st);
void ms_launch_update(const MsSlot *s, const int *active, int prec, int Vh, VecView rs,
                      VecView ps, VecView pm, VecView xm, long stride, long sstride, int cap,
                      int ns, double *partB, hipStream_t st);

class MultiShiftBatch {
 public:
  // sloppy_prec: P32 or P16.  masses: ns >= 1 masses shared by all sources.
  MultiShiftBatch(Links &L, int eo, int capacity, int sloppy_prec, const double *masses, int ns);
  ~MultiShiftBatch();

 private:
  void release();  // frees everything allocated so far

 public:
  MultiShiftBatch(const MultiShiftBatch &) = delete;
  MultiShiftBatch &operator=(const MultiShiftBatch &) = delete;
  // load source b (full FP64 vector; its eo half is used) with a sloppy
  // relative residual target per shift
  void load(int slot, const DevVec &src, double ms_tol, int max_iter);
  void run();  // iterate until every loaded source is done
  // x of (slot, mass index) into out's eo half (FP64, no mass factor)
  void solution(int slot, int k, DevVec &out);
  int iterations(int slot);
  int updates(int slot);  // reliable updates done

 private:
  void enqueue_iteration();
  void enqueue_reliable_update();
  Links &L_;
  Lattice *lat_;
  int eo_, cap_, sp_, ns_, imin_;
  int hp_;  // precision of the sloppy hops: sp_, or P32H (FP32 vectors, FP16 links)
  int nch_;
  std::vector<double> masses_;
  DevBatch b64_, rs_, ps_, ts_;
  DevBatch pm_;  // [ns][cap] sloppy shifted directions (index imin unused)
  DevBatch xm_;  // [ns][cap] FP32 solution increments since the last reliable update
  DevBatch x64_; // [ns][cap] FP64 solutions
  DevBatch t64_, ax64_;  // [cap] FP64 hop temporaries of the reliable update
  int *ru_ = nullptr;    // [cap] reliable update requested
  HopShape ru_shape_;
  MsSlot *slot_ = nullptr, *slot_h_ = nullptr;
  int *active_ = nullptr, *active_h_ = nullptr;
  int *sel_ = nullptr;
  double *partA_ = nullptr, *partB_ = nullptr, *partC_ = nullptr;
  double *coef_ = nullptr, *sums_ = nullptr;
  int *zero_status_ = nullptr;
  CapturedGraph graph_;
  HopShape shape_;
};

}  // namespace qop
