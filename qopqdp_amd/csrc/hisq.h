// HISQ link preparation on the device (hisq.hip, hisq.cpp).
#pragma once
#include "common.h"

namespace qop {

struct U3Args {
  long V = 0;                 // sites; 4 V links are projected
  const double *in[4] = {};   // fat7 links, raw layout [V][3][3][2] per direction
  double *out[4] = {};        // projected links, same layout (may equal in)
  int allow_svd = 1, svd_only = 0;
  double rel_error = 1e-8, abs_error = 1e-8;
  int *counts = nullptr;      // [u3_nwaves][2]: fallbacks, singular links per wave
};
long u3_nwaves(const Lattice *lat);
void launch_u3_project(const U3Args &a, hipStream_t st);
// o = a + s b over n doubles (a null: o = s b)
void launch_axpy_raw(double *o, const double *a, double s, const double *b, long n, hipStream_t st);

}  // namespace qop
