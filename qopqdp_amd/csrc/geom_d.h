// Device helpers for full-lattice fields in the raw site order: raw index i
// runs over the even sites (lexicographic order, x3 fastest, i < Vh) and then
// the odd sites, the order of the raw arrays at the QOP API.
#pragma once
#include "common.h"

namespace qop {

static __device__ __forceinline__ void raw_coords_d(const Geom &g, long i, int x[4]) {
  const long Vh = g.Vh;
  int p = i >= Vh;
  long h = i - p * Vh;
  long l = 2 * h;
  long t = l;
  x[3] = (int)(t % g.L[3]); t /= g.L[3];
  x[2] = (int)(t % g.L[2]); t /= g.L[2];
  x[1] = (int)(t % g.L[1]); t /= g.L[1];
  x[0] = (int)t;
  if (((x[0] + x[1] + x[2] + x[3]) & 1) != p) x[3] += 1;
}

static __device__ __forceinline__ long raw_index_d(const Geom &g, const int x[4]) {
  long l = ((long)(x[0] * g.L[1] + x[1]) * g.L[2] + x[2]) * g.L[3] + x[3];
  int p = (x[0] + x[1] + x[2] + x[3]) & 1;
  return (l >> 1) + (p ? (long)g.Vh : 0);
}

static __device__ __forceinline__ long shift_d(const Geom &g, const int x[4], int mu, int d) {
  int y[4] = {x[0], x[1], x[2], x[3]};
  y[mu] = ((y[mu] + d) % g.L[mu] + g.L[mu]) % g.L[mu];
  return raw_index_d(g, y);
}

}  // namespace qop
