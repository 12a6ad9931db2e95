// Host-side field operations on device vectors (the data-parallel layer the
// solvers are written in).  Subsets follow QOP_evenodd_t: 0 even, 1 odd,
// 2 both (lib/asqtad_dslash_p.c, include/qop_internal.h:64-65).
#pragma once
#include "common.h"
#include "kernels.h"

namespace qop {

struct Cplx {
  double re, im;
};

inline int oppsub(int eo) { return (4 - eo) % 3; }

// dst(out_parity) = alpha * hop(src) + beta * x(out_parity); optional FP64
// reduction (RED_OUT2 / RED_X2) written to dev_result[0] (device double).
// Link arguments of a one-RHS hop onto out_parity: single-stored links when
// the QOP style option is even (lib/asqtad_dslash_p.c:8, dblstore_style) and
// the precision is FP64/FP32, the double-stored set otherwise.
void set_hop_links(HopArgs &a, Links &L, int prec, int out_parity);

void hop(Links &L, int prec, DevVec &dst, int out_parity, const DevVec &src,
         double alpha, double beta, const DevVec *x, int red, double *dev_result,
         hipStream_t st, const int *status = nullptr);

// QOP_asqtad_dslash semantics on device vectors
// (lib/asqtad_dslash_p.c:1048-1092).
void dslash(Links &L, double mass, DevVec &out, const DevVec &in, int eo_out,
            int eo_in);

double norm2(const DevVec &v, int eo);
Cplx cdot(const DevVec &a, const DevVec &b, int eo);  // sum conj(a) b
double redot(const DevVec &a, const DevVec &b, int eo);
// Fermilab relative residual (lib/invert_utilities_p.c:7-40)
double relnorm2(const DevVec &r, const DevVec &x, int eo);

// out = sum_i c_i v_i on subset eo (inputs may alias out)
void lincomb(DevVec &out, int eo, int nin, const Cplx *c, const DevVec *const *v,
             hipStream_t st = nullptr);
void copy(DevVec &d, const DevVec &s, int eo);
void zero(DevVec &d, int eo);
void axpy(DevVec &d, double a, const DevVec &x, const DevVec &y, int eo);  // d = a x + y
void caxpy(DevVec &d, Cplx a, const DevVec &x, int eo);                    // d += a x
void scal(DevVec &d, Cplx a, const DevVec &x, int eo);                     // d = a x

// Device-vector pool for solver temporaries (same lattice and precision).
struct TmpVec {
  DevVec v;
  TmpVec(Lattice *lat, int prec) { devvec_alloc(v, lat, prec); }
  ~TmpVec() { devvec_free(v); }
  TmpVec(const TmpVec &) = delete;
  TmpVec &operator=(const TmpVec &) = delete;
};

}  // namespace qop
