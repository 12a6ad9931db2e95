// Kernel argument blocks and launchers (see kernels.hip).
#pragma once
#include "common.h"

namespace qop {

// Hop reductions per RHS (FP64): RED_OUT2 = |out|^2, RED_X2 = |x|^2.
// RED_CG: fused CG update in the epilogue (dst = r is updated in place with
// r -= alpha (beta x + alpha_hop hop), xs += alpha x, reduction |r|^2).
enum { RED_NONE = 0, RED_OUT2 = 1, RED_X2 = 2, RED_CG = 3 };
int hop_partial_blocks(const Lattice *lat);

// Hop: dst[b] = alpha_b * sum_k W_k src[b](x+d_k) + beta_b * x[b], for the
// output parity; slot b of a view is at data + b*stride complex elements and
// norm + b*nstride floats.
struct HopArgs {
  Geom g;
  int parity = 0;
  const void *links = nullptr;   // links for this output parity: [nl][9][Vh]
  const float *lscale = nullptr; // FP16: [nl][Vh]
  VecView src{}, dst{}, x{};
  long sstride = 0, dstride = 0, xstride = 0, nstride = 0;
  int n = 1;
  const double *coef = nullptr;  // device [n][2] (alpha, beta) or null
  double alpha = 1.0, beta = 0.0;
  double *partial = nullptr;     // [n][nchunk]
  const int *active = nullptr;   // device [n] or null
  const int *slots = nullptr;    // device [1 + n] or null: count k, then the slots that launch
                                 // RHS 0..k-1 work on (FP64 masked hops: k flagged slots take
                                 // ceil(k / RHS per thread) link reads)
  const int *status = nullptr;   // device flag: kernel is a no-op when nonzero
  int remap = 1;
  int nt = 0;                    // non-temporal link loads
  int rhs_inner = 0;             // RHS groups innermost in the block order
  int tile = 0;                  // x0-tiled site order: sites per tile of an x0 slab (0: plain)
  int stage = 0;                 // LDS staging: the block's own source-parity sites are
                                 // loaded once into LDS and the neighbours inside that
                                 // set are read from there; 1: the block's rows (x3 and
                                 // most x2 hops), 2: plane-blocked order (also the x0
                                 // hops inside the block's planes)
  // single-stored links: `links` holds the forward links of the output
  // parity, links_bck those of the source parity ([nl/2][9][Vh] each)
  int single = 0;
  const void *links_bck = nullptr;
  VecView xs{};                  // RED_CG: FP32 solution increments (stride dstride)
  const double *cg_alpha = nullptr;  // RED_CG: alpha of slot b at cg_alpha[b*cg_stride]
  int cg_stride = 0;
  int ngroups = 1;               // set by launch_hop_args
  int buf = 1;                   // set by launch_hop_args: FP32 / FP16 buffer loads
  int plain = 0;                 // FP32 / FP16: one RHS per thread with plain loads (the
                                 // instance lattices beyond the buffer range take)
};
void launch_hop_args(HopArgs a, int prec, int nl, int nb, int block, int red,
                     bool hasx, hipStream_t st);

// out[b] = sum_i c_i v_i[b] (complex c_i; coef device [n][3][2] or host cr/ci)
struct LincombArgs {
  int Vh = 0;
  int n = 1;
  VecView out{};
  VecView v[3]{};
  long ostride = 0, stride[3] = {0, 0, 0}, nstride = 0;
  double cr[3] = {0, 0, 0}, ci[3] = {0, 0, 0};
  const double *coef = nullptr;
  int coef_neg = 0;               // bit i: negate device coefficient i
  double *partial = nullptr;      // |out|^2 per chunk when requested
  const int *active = nullptr;
  const int *status = nullptr;
};
void launch_lincomb(const LincombArgs &a, int po, const int *pin, int nin,
                    bool red, hipStream_t st);

struct DotArgs {
  int Vh = 0;
  int n = 1;
  VecView u{}, v{};
  long ustride = 0, vstride = 0, nstride = 0;
  double *partial = nullptr;  // mode 0: [n][nchunk]; mode 1: [2n][nchunk]
  const int *active = nullptr;
  const int *status = nullptr;
};
void launch_dot(const DotArgs &a, int pu, int pv, int mode, hipStream_t st);

void launch_finalize(const double *partial, int m, int n, double *result,
                     hipStream_t st, const int *status);

// One-direction transport onto out_parity (QOP_asqtad_dslash_dir):
// out(x) += wf W_a(x) in(x+d_a) + wl W_b(x) in(x+d_b) with the double-stored
// links of direction dir (a, b = fat and long, forward for fb > 0 and
// backward otherwise; bsign multiplies the backward terms).  Only the fat
// term without Naik links.
void launch_hop_dir(const Lattice *lat, const LinkSet &ls, int prec, int out_parity, int dir, int fb,
                    double wf, double wl, double bsign, VecView out, VecView in, hipStream_t st);

// streaming copy of `bytes` (multiple of 16)
void launch_copy(const void *in, void *out, size_t bytes, hipStream_t st);

}  // namespace qop
