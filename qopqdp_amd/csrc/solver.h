// Solver state shared between host drivers and device scalar kernels.
#pragma once
#include "common.h"
#include "ops.h"

namespace qop {

constexpr int CGMS_MAXSHIFT = 64;

// Multi-shift CG state (lib/invert_cg_p.c:130-402).  It lives in device
// memory; the per-iteration scalar recurrences run on the GPU so the host
// does not synchronise every iteration.  status: 0 running, 1 the host must
// run the solution fold / true-residual check, 2 finished.
struct CgmsState {
  int status;
  int iteration;
  int ns, imin, irsq, irel;
  int restart, max_iter;
  int nextRsqCheck, nextRelCheck;
  int break_reason;  // 1 pkp<=0, 2 restart/max_iter
  int pad;
  double rsq, oldrsq, pkp, insq, rsqstop, relstop, trueRsqFac, trueRelFac, relnorm2;
  double m2;  // base mass squared of the linear operator
  double shifts[CGMS_MAXSHIFT], a[CGMS_MAXSHIFT], b[CGMS_MAXSHIFT],
      bo[CGMS_MAXSHIFT], z[CGMS_MAXSHIFT], zo[CGMS_MAXSHIFT], zn[CGMS_MAXSHIFT];
};

struct CgmsDev {
  CgmsState *st;       // device state
  double *partA, *partB, *partC, *partD;  // reductions per 64-site chunk
  VecView *xs, *pms;   // device arrays [ns] of half views (x_i, p_i)
};

// Launchers (cgms.hip)
void cgms_launch_scalarA(CgmsDev d, int nch, hipStream_t st);
void cgms_launch_update(CgmsDev d, int prec, int Vh, VecView r, VecView Mp, int ns,
                        hipStream_t st);
void cgms_launch_scalarB(CgmsDev d, int nch, int with_rel, int Vh, hipStream_t st);
void cgms_launch_pupdate(CgmsDev d, int prec, int Vh, VecView p, VecView r, int ns,
                         hipStream_t st);
void cgms_launch_relnorm(CgmsDev d, int prec, int Vh, VecView r, VecView x,
                         hipStream_t st);

// QOP option block (lib/asqtad_invert.c:8-112) plus this library's options.
struct Options {
  int style = 1, nsvec = 8, nvec = 8, cgtype = 0;
  int eigcg_nev = 10, eigcg_m = 100, eigcg_numax = 200;
  // library extensions (tags "qmx_solver", "qmx_prec", ...)
  int solver = 0;        // 0: reference algorithm; 1: reliable-update mixed precision
  int inner_prec = P32;  // inner precision of the mixed solver
  double delta = 0.1;    // reliable-update residual drop
  int chunk = 16;        // iterations enqueued between host status checks
  int use_graph = 1;     // capture the iteration chunk in a hipGraph
  double ru_join = 0.0;  // batched CG: a slot whose residual dropped by this factor since its
                         // last reliable update joins another slot's update (0: never)
  int ms_prec = P32;     // vector precision of the batched multi-shift stage
  int ms_link16 = 0;     // FP32 multi-shift stage: hop with the FP16 links (+ scales)
  int verbose = 0;
};
Options &options();

struct InvArg {
  double mixed_rsq;
  int max_iter, restart, max_restarts, evenodd;
};
struct ResArg {
  double rsqmin, final_rsq, relmin, final_rel;
  int final_iter, final_restart;
};

// Reference-algorithm drivers (solver_ref.cpp); vectors FP64 unless noted.
void invert_cgms(Links &L, int prec, double mass, int eo, InvArg *ia, ResArg **ra,
                 const double *shifts, int ns, DevVec **out, const DevVec &in);
void invert_cg(Links &L, int prec, double mass, int eo, InvArg *ia, ResArg *ra,
               DevVec &out, const DevVec &in);
void solve_multi(Links &L, InvArg *ia, ResArg **ra, const double *masses,
                 DevVec **out, DevVec *const *in, int nsolve, double *final_flop,
                 double *final_sec);
void invert(Links &L, InvArg *ia, ResArg *ra, double mass, DevVec &out,
            const DevVec &in, double *final_flop, double *final_sec);
void invert_multi(Links &L, InvArg *ia, ResArg **ra, const double *masses,
                  const int *nmass, DevVec **out, DevVec *const *in, int nsrc,
                  double *final_flop, double *final_sec);

// <p, A p> with A = m^2 - D_eo D_oe on subset eo (lib/asqtad_dslash_p.c:1121-1138)
double ddagd_norm2(Links &L, int prec, double mass, DevVec &out, const DevVec &in,
                   int eo, DevVec &tmp);

double wall_time();

}  // namespace qop
