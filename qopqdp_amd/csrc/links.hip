// Link preparation kernels on gfx950:
//  * double storage of user-normalised fat/long links with the x1/2 factor
//    (lib/asqtad_dslash_p.c:79-133, 152-223) into the hop's SoA layout;
//  * precision copies (FP64 -> FP32, FP64 -> FP16 + per-link float scale);
//  * extraction back to raw (x2, lib/asqtad_dslash_p.c:915-938);
//  * staggered phases and boundary phases (lib/common_p.c:464-526,
//    lib/asqtad_dslash_p.c:1385-1482);
//  * asqtad fattening (lib/asqtad_dslash_p.c:391-633) of raw-layout links,
//    computed by fatten_dev (force.hip) on site-innermost copies.
#include "common.h"
#include "geom_d.h"
#include "kernels.h"

namespace qop {

// ---------------------------------------------------------------------------
// double store: out[p][k][e][h]
__global__ void dblstore_kernel(Geom g, const double *const *fat,
                                const double *const *lng, int nl, double2 *out) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long V = 2L * g.Vh;
  if (i >= V) return;
  int x[4];
  raw_coords_d(g, i, x);
  const int p = i >= g.Vh;
  const long h = i - p * g.Vh;
  const int per = nl / 4;
  for (int mu = 0; mu < 4; mu++) {
    for (int j = 0; j < per; j++) {
      const bool fwd = per == 4 ? (j < 2) : (j == 0);
      const bool isl = per == 4 && (j & 1);
      const double *src = isl ? lng[mu] : fat[mu];
      const int d = isl ? 3 : 1;
      const int k = per * mu + j;
      double2 *o = out + ((long)p * nl + k) * 9 * g.Vh + h;
      if (fwd) {
        const double *m = src + 18 * i;
        for (int e = 0; e < 9; e++) o[(long)e * g.Vh] = make_double2(0.5 * m[2 * e], 0.5 * m[2 * e + 1]);
      } else {
        const double *m = src + 18 * shift_d(g, x, mu, -d);
        // -(1/2 m)^dagger: element (r,c) = -conj(m[c][r]) / 2
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++)
            o[(long)(3 * r + c) * g.Vh] =
                make_double2(-(0.5 * m[6 * c + 2 * r]), 0.5 * m[6 * c + 2 * r + 1]);
      }
    }
  }
}

void build_links64(Links &L, const double *fat_dev[4], const double *lng_dev[4],
                   hipStream_t st) {
  const Lattice *lat = L.lat;
  L.nl = lng_dev ? 16 : 8;
  LinkSet &s = L.l64;
  linkset_free(s);
  L.drop_derived();
  DevBuf<const double *> pbuf(8 * sizeof(double *));
  const double **ptrs = pbuf.p;
  dev_alloc(&s.data, (size_t)2 * L.nl * 9 * lat->Vh * sizeof(double2));
  s.prec = P64;
  s.nl = L.nl;
  const double *hp[8];
  for (int i = 0; i < 4; i++) { hp[i] = fat_dev[i]; hp[4 + i] = lng_dev ? lng_dev[i] : nullptr; }
  QOP_CHECK_HIP(hipMemcpyAsync(ptrs, hp, sizeof hp, hipMemcpyHostToDevice, st));
  dim3 g((lat->V + 255) / 256), b(256);
  hipLaunchKernelGGL(dblstore_kernel, g, b, 0, st, lat->g, ptrs, ptrs + 4, s.nl, (double2 *)s.data);
  QOP_CHECK_HIP(hipGetLastError());
  QOP_CHECK_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// precision copies: one thread per (p,k,h) link
__global__ void to32_kernel(const double2 *in, float2 *out, long nlinks, long Vh) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nlinks) return;
  long pk = t / Vh, h = t - pk * Vh;
  const double2 *a = in + pk * 9 * Vh + h;
  float2 *o = out + pk * 9 * Vh + h;
  for (int e = 0; e < 9; e++) {
    double2 v = a[e * Vh];
    o[e * Vh] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ void to16_kernel(const double2 *in, __half2 *out, float *scale,
                            long nlinks, long Vh) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nlinks) return;
  long pk = t / Vh, h = t - pk * Vh;
  const double2 *a = in + pk * 9 * Vh + h;
  __half2 *o = out + pk * 9 * Vh + h;
  double m = 0;
  for (int e = 0; e < 9; e++) {
    double2 v = a[e * Vh];
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
  float s = (float)m;
  float inv = s > 0 ? 1.0f / s : 0.0f;
  for (int e = 0; e < 9; e++) {
    double2 v = a[e * Vh];
    // imaginary part negated: the hop forms Re(w v) = wr vr - wi vi as one
    // two-term dot product of the stored word and the vector word
    o[e * Vh] = __floats2half2_rn((float)v.x * inv, -((float)v.y * inv));
  }
  scale[t] = s;
}

void convert_links(const Lattice *lat, const LinkSet &src, LinkSet &dst, int prec,
                   hipStream_t st) {
  QOP_REQUIRE(src.prec == P64, "convert_links: source must be FP64");
  QOP_REQUIRE(prec == P32 || prec == P16, "convert_links: bad precision");
  linkset_free(dst);
  long nlinks = 2L * src.nl * lat->Vh;
  // allocate both parts before the set is marked built, so that a failed
  // allocation leaves it empty
  DevBuf<void> data((size_t)nlinks * 9 * prec_cbytes(prec));
  DevBuf<float> scale;
  if (prec == P16) dev_alloc(&scale.p, (size_t)nlinks * sizeof(float));
  dst.data = data.release();
  dst.scale = scale.release();
  dst.prec = prec;
  dst.nl = src.nl;
  dim3 g((nlinks + 255) / 256), b(256);
  if (prec == P32) {
    hipLaunchKernelGGL(to32_kernel, g, b, 0, st, (const double2 *)src.data, (float2 *)dst.data, nlinks, lat->Vh);
  } else if (prec == P16) {
    hipLaunchKernelGGL(to16_kernel, g, b, 0, st, (const double2 *)src.data, (__half2 *)dst.data, dst.scale, nlinks, lat->Vh);
  } else {
    fatal(__FILE__, __LINE__, "convert_links: bad precision");
  }
  QOP_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// extract: raw fat/long (user normalisation) from the forward FP64 links
__global__ void extract_kernel(Geom g, const double2 *links, int nl, double *const *fat,
                               double *const *lng) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long V = 2L * g.Vh;
  if (i >= V) return;
  const int p = i >= g.Vh;
  const long h = i - p * g.Vh;
  const int per = nl / 4;
  for (int mu = 0; mu < 4; mu++) {
    for (int j = 0; j < (per == 4 ? 2 : 1); j++) {
      double *dst = j == 0 ? fat[mu] : lng[mu];
      if (!dst) continue;
      const int k = per * mu + j;
      const double2 *w = links + ((long)p * nl + k) * 9 * g.Vh + h;
      for (int e = 0; e < 9; e++) {
        double2 v = w[(long)e * g.Vh];
        dst[18 * i + 2 * e] = 2.0 * v.x;
        dst[18 * i + 2 * e + 1] = 2.0 * v.y;
      }
    }
  }
}

void extract_links64(const Links &L, double *fat_dev[4], double *lng_dev[4], hipStream_t st) {
  const Lattice *lat = L.lat;
  DevBuf<double *> pbuf(8 * sizeof(double *));
  double **ptrs = pbuf.p;
  double *hp[8];
  for (int i = 0; i < 4; i++) { hp[i] = fat_dev[i]; hp[4 + i] = lng_dev ? lng_dev[i] : nullptr; }
  QOP_CHECK_HIP(hipMemcpyAsync(ptrs, hp, sizeof hp, hipMemcpyHostToDevice, st));
  dim3 g((lat->V + 255) / 256), b(256);
  hipLaunchKernelGGL(extract_kernel, g, b, 0, st, lat->g, (const double2 *)L.l64.data, L.nl, ptrs, ptrs + 4);
  QOP_CHECK_HIP(hipGetLastError());
  QOP_CHECK_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// phases on one direction's raw matrices.  bcdir: apply bc on the last
// `nslice` slices of direction mu (relative to r0); sign bits use (x - r0) when
// sign_rel, else absolute x (rephase_G uses absolute coordinates).
__global__ void rephase_kernel(Geom g, double *U, int mu, int r00, int r01, int r02,
                               int r03, double bcre, double bcim, int bcdir,
                               int signbits, int nslice, int sign_rel) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long V = 2L * g.Vh;
  if (i >= V) return;
  int x[4];
  raw_coords_d(g, i, x);
  const int r0[4] = {r00, r01, r02, r03};
  double *m = U + 18 * i;
  if (bcdir) {
    int rs = (x[mu] + g.L[mu] - r0[mu]) % g.L[mu];
    if (rs >= g.L[mu] - nslice) {
      for (int e = 0; e < 9; e++) {
        double a = m[2 * e], b = m[2 * e + 1];
        m[2 * e] = bcre * a - bcim * b;
        m[2 * e + 1] = bcre * b + bcim * a;
      }
    }
  }
  if (signbits) {
    int s = 0;
    for (int j = 0; j < 4; j++) {
      int c = sign_rel ? (x[j] + g.L[j] - r0[j]) % g.L[j] : x[j];
      s += ((signbits >> j) & 1) * c;
    }
    if (s & 1)
      for (int e = 0; e < 18; e++) m[e] = -m[e];
  }
}

void launch_rephase(const Lattice *lat, double *U_dev, int mu, const int r0[4],
                    double bcre, double bcim, int bcdir, int signbits, int nslice,
                    int sign_rel, hipStream_t st) {
  dim3 g((lat->V + 255) / 256), b(256);
  hipLaunchKernelGGL(rephase_kernel, g, b, 0, st, lat->g, U_dev, mu, r0[0], r0[1], r0[2],
                     r0[3], bcre, bcim, bcdir, signbits, nslice, sign_rel);
  QOP_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// fattening (compute_gen_staple / make_imp_links): site-innermost kernels
// shared with the force (force.hip, fatten_dev)
void launch_fatten(const Lattice *lat, const double *gf[4], const double *gl[4],
                   const double c[6], double *fat[4], double *lng[4], hipStream_t st) {
  fatten_dev(lat, gf, gl, c, fat, lng, st);
}



// ---------------------------------------------------------------------------
// one-direction transport (lib/asqtad_dslash_p.c:1140-1191, 1338-1377): one
// thread per output site; not on the solvers' path
template <typename C>
__global__ void hop_dir_kernel(Geom g, int parity, const C *__restrict__ links, int nl, int dir, int fb,
                               double wf, double wl, double bsign, C *out, const C *__restrict__ in) {
  const long h = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (h >= g.Vh) return;
  const long Vh = g.Vh;
  int x[4];
  raw_coords_d(g, h + parity * Vh, x);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int nterm = nl == 16 ? 2 : 1;
  for (int j = 0; j < nterm; j++) {
    const int k = nl == 16 ? 4 * dir + (fb > 0 ? 0 : 2) + j : 2 * dir + (fb > 0 ? 0 : 1);
    const int d = (j ? 3 : 1) * (fb > 0 ? 1 : -1);
    const long nb = shift_d(g, x, dir, d) - (long)(1 - parity) * Vh;
    const double w = (j ? wl : wf) * (fb > 0 ? 1.0 : bsign);
    const C *W = links + (long)k * 9 * Vh + h;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        const C m = W[(long)(3 * r + c) * Vh];
        const C v = in[(long)c * Vh + nb];
        acc[2 * r] += w * ((double)m.x * v.x - (double)m.y * v.y);
        acc[2 * r + 1] += w * ((double)m.x * v.y + (double)m.y * v.x);
      }
  }
  for (int r = 0; r < 3; r++) {
    C o = out[(long)r * Vh + h];
    o.x = o.x + acc[2 * r];
    o.y = o.y + acc[2 * r + 1];
    out[(long)r * Vh + h] = o;
  }
}

void launch_hop_dir(const Lattice *lat, const LinkSet &ls, int prec, int out_parity, int dir, int fb,
                    double wf, double wl, double bsign, VecView out, VecView in, hipStream_t st) {
  QOP_REQUIRE(prec == P64 || prec == P32, "dslash_dir: FP64 or FP32 vectors");
  QOP_REQUIRE(ls.prec == prec && (ls.nl == 16 || ls.nl == 8), "dslash_dir: link set");
  dim3 g((lat->Vh + 255) / 256), b(256);
  if (prec == P64)
    hipLaunchKernelGGL(hop_dir_kernel<double2>, g, b, 0, st, lat->g, out_parity,
                       (const double2 *)ls.parity_data(lat, out_parity), ls.nl, dir, fb, wf, wl, bsign,
                       (double2 *)out.data, (const double2 *)in.data);
  else
    hipLaunchKernelGGL(hop_dir_kernel<float2>, g, b, 0, st, lat->g, out_parity,
                       (const float2 *)ls.parity_data(lat, out_parity), ls.nl, dir, fb, wf, wl, bsign,
                       (float2 *)out.data, (const float2 *)in.data);
  QOP_CHECK_HIP(hipGetLastError());
}

}  // namespace qop
