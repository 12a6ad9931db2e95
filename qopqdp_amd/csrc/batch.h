// Batched mixed-precision CG with reliable updates.
//
// Each slot b solves (m_b^2 - D_eo D_oe) x_b = src_b on one parity of one
// lattice, sharing the links of the batch.  The Krylov iteration runs in a
// sloppy precision (FP32, or FP16 with per-site float norms) with FP64
// reductions and scalars; whenever the iterated residual has dropped by
// `delta` since the last reliable update (or appears converged, or hits the
// iteration limit) the slot's FP64 solution is updated, its FP64 true
// residual b - A x is recomputed and the search direction is re-orthogonalised
// against it.  A slot stops when its FP64 true relative residual reaches its
// tolerance.
//
// Every per-slot quantity (hop, reductions in a fixed order, updates) is
// computed the same way whatever the other slots hold, so a slot's result
// does not depend on what it is batched with.  Slots can be loaded and
// released between iteration chunks (continuous batching).
#pragma once
#include "solver.h"

namespace qop {

struct BsRhs {
  int need_update;  // reliable update requested by the scalar kernel
  int iter;         // sloppy iterations done
  int max_iter;
  int reason;       // 0 running, 1 converged, 2 iteration limit, 3 breakdown
  int nupdate;      // reliable updates done
  int pad;
  double m2;        // mass^2 of the slot's operator
  double r2;        // iterated residual norm^2
  double r2_old;    // iterated residual norm^2 of the previous iteration
  double r2_rel;    // true residual norm^2 at the last reliable update
  double pAp, alpha, beta;
  double stop2;     // tol^2 |src|^2
  double b2;        // |src|^2
  double delta2;    // reliable-update threshold (delta^2)
  double join2;     // joins another slot's update below join2 * r2_rel (0: never)
};

struct BsGlobal {
  int status;   // 0 running, 1 host attention needed (read by every kernel)
  int pending;  // set by slots during an iteration, committed at its end
};

// Kernels (batch.hip)
void bs_launch_scalar1(BsRhs *rhs, const int *active, const int *status,
                       const double *partA, const double *partB, int nch, int n,
                       int *ru_cont, hipStream_t st);
void bs_launch_update(BsRhs *rhs, const int *active, const int *status, int prec,
                      int Vh, VecView rs, VecView Aps, VecView ps, VecView xs,
                      long stride, double *partC, int n, hipStream_t st);
void bs_launch_scalar2(BsRhs *rhs, const int *active, BsGlobal *glob,
                       const double *partC, int nch, int n, int *ru, hipStream_t st);
// Device reliable update, stage 0: x += xs; 1: residual, |r|^2 and <rs,ps>;
// 2: per-slot decisions; 3: search-direction fix.
// list: compacted flagged slots (bs_launch_compact of ru), read by the
// vector steps 0, 1 and 3
void bs_launch_ru(BsRhs *rhs, int *active, int *ru, int *ru_cont, const int *list, double *lcoef,
                  int prec, int Vh, VecView x64, VecView xs, VecView b64, VecView ax64, VecView rs,
                  VecView ps, long stride, double *pr2, double *pre, double *pim, int nch,
                  int n, int stage, double *ppB, hipStream_t st);
// A slot whose residual has dropped to join2 * r2_rel joins the reliable
// update of this iteration when another slot is flagged for one.
void bs_launch_ru_join(BsRhs *rhs, const int *active, int *ru, int n, hipStream_t st);
// out[0] = number k of nonzero flags[0..n), out[1..k] their indices (device)
void bs_launch_compact(const int *flags, int n, int *out, hipStream_t st);
void bs_launch_commit(BsGlobal *glob, hipStream_t st);
void bs_launch_pupdate(const BsRhs *rhs, const int *active, const int *status,
                       int prec, int Vh, VecView rs, VecView ps, long stride, int n,
                       int skip_flagged, double *partB, hipStream_t st);
void bs_launch_sum(const double *part, int nch, int n, const int *sel,
                   double *out, hipStream_t st);

// Tuned hop launch shape for (lattice, precision, nl, nrhs), timed on first use
// and cached (batch.cpp).
struct HopShape {
  // plain: force the one-RHS-per-thread plain-load instance (FP32 / FP16)
  int nb = 1, block = 256, remap = 1, nt = 0, rhs_inner = 0, tile = 0, stage = 0, plain = 0;
  void apply(HopArgs &a) const {
    a.remap = remap;
    a.nt = nt;
    a.rhs_inner = rhs_inner;
    a.tile = tile;
    a.stage = stage;
    a.plain = plain;
  }
};
struct SlotResult {
  int iter = 0, nupdate = 0, reason = 0;
  double true_rsq = 0;  // |src - A x|^2 / |src|^2 in FP64
};

class BatchSolver {
 public:
  // st: stream the solver's work runs on (nullptr: the library stream)
  BatchSolver(Links &L, int eo, int capacity, int sloppy_prec, double delta,
              hipStream_t st = nullptr);
  ~BatchSolver();

 private:
  void release();  // frees everything allocated so far

 public:
  BatchSolver(const BatchSolver &) = delete;
  BatchSolver &operator=(const BatchSolver &) = delete;

  int capacity() const { return cap_; }
  int free_slot() const;
  bool slot_free(int slot) const { return !loaded_[slot]; }
  int nactive() const;
  // Load a system into a free slot.  src is a full FP64 vector (its `eo`
  // half is used); x0 (optional) is an initial guess for A^-1 src.
  void load(int slot, const DevVec &src, const DevVec *x0, double mass, double tol,
            int max_iter);
  // several slots at once (one masked FP64 operator application for all
  // initial residuals); the same per-slot state as load()
  struct LoadReq {
    int slot;
    const DevVec *src;
    const DevVec *x0;
    double mass, tol;
    int max_iter;
  };
  void load_many(const LoadReq *req, int nreq);
  // Run one chunk of iterations and service reliable updates.  Slots that
  // finished are appended to `done`.
  void step(std::vector<int> &done);
  // step() in two halves: launch_chunk() enqueues one chunk and returns at
  // once; after chunk_done() reports true, collect() lists finished slots.
  // Nothing else may be called on the solver in between.
  void launch_chunk();
  bool chunk_done();
  void collect(std::vector<int> &done);
  bool in_flight() const { return inflight_; }
  hipStream_t stream_of() const { return st_; }
  int eo() const { return eo_; }
  Lattice *lattice() const { return lat_; }
  // Write mass * x (the m/A normalisation) or x into out's eo half.
  void result(int slot, DevVec &out, bool times_mass, SlotResult *res);
  void release(int slot);
  // Run until every loaded slot is done.
  void run_all(std::vector<int> &done);
  double mass(int slot) const { return mass_[slot]; }

 private:
  void enqueue_iteration();
  void build_graph();
  // ax64 = A x64 on the slots set in mask (listed: slots_ already holds
  // their compacted list)
  void apply_A64(const int *mask, bool listed = false);

  Links &L_;
  Lattice *lat_;
  int eo_, cap_, sp_;
  double delta_;
  double join_;  // reliable-update joining threshold (residual drop), 0: off
  hipStream_t st_;
  hipEvent_t ev_ = nullptr;
  bool inflight_ = false;
  int nch_;
  // FP64 per slot
  DevBatch x64_, r64_, b64_, ax64_, t64_;
  // sloppy
  DevBatch rs_, ps_, ts_, xs_;
  BsRhs *rhs_ = nullptr, *rhs_h_ = nullptr;
  int *active_ = nullptr, *active_h_ = nullptr;
  int *sel_ = nullptr, *sel_h_ = nullptr;
  int *ru_ = nullptr, *ru_cont_ = nullptr;  // device reliable-update masks
  int *slots_ = nullptr;  // device [1 + cap]: compacted slot list of the masked FP64 hops
  BsGlobal *glob_ = nullptr, *glob_h_ = nullptr;
  double *partA_ = nullptr, *partB_ = nullptr, *partC_ = nullptr;
  double *coef_ = nullptr;      // [cap][2] hop coefficients (-1, m2)
  double *lcoef_ = nullptr;     // [cap][3][2] lincomb coefficients
  double *sums_ = nullptr, *sums_h_ = nullptr;
  double *hc_h_ = nullptr;      // pinned staging of the loaded slots' hop coefficients
  std::vector<int> loaded_;
  std::vector<double> mass_;
  CapturedGraph graph_;
  int graph_chunk_ = 0;
  HopShape shape_;     // tuned launch shape of the first sloppy hop
  HopShape shape2_;    // tuned launch shape of the second (fused CG update)
  HopShape ru_shape_;  // tuned FP64 hop launch shape of the reliable updates
};

// red: the epilogue to tune with (RED_NONE, RED_OUT2, or RED_CG with x and
// xs given; dst and xs keep their values, alpha is 0).
HopShape tune_hop(Links &L, int prec, int nrhs, const DevBatch &src, DevBatch &dst,
                  int eo_out, int red = RED_NONE, const DevBatch *x = nullptr,
                  const DevBatch *xs = nullptr);
void set_hop_shape_override(const HopShape *s);  // null: tune again
// The shape tune_hop chose for (links, prec, nrhs, red), if it has run (or the
// forced shape): no timing and no allocation.
bool cached_hop_shape(Links &L, int prec, int nrhs, int red, HopShape *out);

}  // namespace qop
