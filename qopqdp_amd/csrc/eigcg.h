// eigCG deflation state and solver (lib/invert_eigcg_p.c:618-1068), the
// inner solver of the EVENODD invert when the option cg = 1
// (lib/asqtad_invert_p.c:683-687, 767-769).  The state belongs to a links
// object, as the reference keeps it in QOP_FermionLinksAsqtad.
#pragma once
#include <vector>

#include "common.h"
#include "solver.h"

namespace qop {

struct EigcgState {
  int numax = 0, m = 0, nev = 0;  // options eigcg_l, eigcg_m, eigcg_nev
  int nu = 0, nv = 0, nn = 0, addvecs = 1, cap = 0, eo = -1;
  Lattice *lat = nullptr;
  double2 *store = nullptr;  // 2 cap parity halves: u[0..cap), rotation output [cap..2cap)
  std::vector<double> l;     // Ritz values of u
  ~EigcgState();
};

// out = A^-1 in on subset eo with eigCG, A = m^2 - D_eo D_oe (FP64).
void invert_eigcg(Links &L, double mass, int eo, InvArg *ia, ResArg *ra, DevVec &out,
                  const DevVec &in, EigcgState &eg);

// y_j = sum_i R(i, j) x_i, j < n2, over len complex elements (eigcg.hip);
// x, y: device arrays of n1 / n2 device pointers; R: device, complex,
// row-major with leading dimension ld.
void launch_rotate(const double2 *const *x, double2 *const *y, const double2 *R, int n1, int n2,
                   int ld, long len, hipStream_t st);

// G(i, j) = sum_k conj(X_i[k]) Y_j[k] over K complex elements, vectors at
// stride ld (eigcg.hip); partial: nblk * n1 * n2 complex scratch; G: device
// n1 x n2 row-major.  Needs n1 n2 <= gram_max_pairs(), n1 + n2 <= gram_max_rows().
void launch_gram(const double2 *X, const double2 *Y, long ld, int n1, int n2, long K, double2 *partial,
                 int nblk, double2 *G, hipStream_t st);
int gram_max_pairs();
int gram_max_rows();

}  // namespace qop
