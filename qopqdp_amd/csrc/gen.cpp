// Seeded synthetic inputs (host side): SU(3) gauge links and Gaussian
// colour-vector sources in the QOP raw layout.  This plays the role of the
// reference's examples/test_common.c input preparation; it is workload
// generation, not part of the solve.
//
// Random numbers: splitmix64 counter streams, one per (site, direction) or
// site, so fields are identical for any thread count; Gaussians by
// Box-Muller.
//
// Gauge recipe: U = reunit(exp(i*eps*H)), H = A - tr(A)/3, A = (G+G^+)/2,
// G a complex Gaussian 3x3 matrix; reunit = row Gram-Schmidt followed by a
// phase fix of the last column so det U = 1 (examples/test_common.c:138-168).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/qop_mi355x.h"

namespace {

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed, uint64_t stream) {
    s = seed * 0x9E3779B97F4A7C15ull ^ (stream + 1) * 0xD1B54A32D192ED03ull;
    next();
    next();
  }
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  void gauss2(double &a, double &b) {
    double u1 = uniform(), u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    a = r * std::cos(2 * M_PI * u2);
    b = r * std::sin(2 * M_PI * u2);
  }
};

using M3 = double[18];

void mmul(double *c, const double *a, const double *b) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += a[6 * i + 2 * k] * b[6 * k + 2 * j] - a[6 * i + 2 * k + 1] * b[6 * k + 2 * j + 1];
        im += a[6 * i + 2 * k] * b[6 * k + 2 * j + 1] + a[6 * i + 2 * k + 1] * b[6 * k + 2 * j];
      }
      t[6 * i + 2 * j] = re;
      t[6 * i + 2 * j + 1] = im;
    }
  memcpy(c, t, sizeof t);
}

// exp(X) by scaling and squaring with a 16-term Taylor series.
void mexp(double *out, const double *X) {
  double n = 0;
  for (int e = 0; e < 18; e++) n += X[e] * X[e];
  n = std::sqrt(n);
  int s = 0;
  while (n > 0.25) { n *= 0.5; s++; }
  double Y[18], term[18], acc[18];
  double f = std::ldexp(1.0, -s);
  for (int e = 0; e < 18; e++) Y[e] = X[e] * f;
  memset(acc, 0, sizeof acc);
  memset(term, 0, sizeof term);
  for (int i = 0; i < 3; i++) { acc[8 * i] = 1; term[8 * i] = 1; }
  for (int k = 1; k <= 16; k++) {
    mmul(term, term, Y);
    for (int e = 0; e < 18; e++) term[e] /= k;
    for (int e = 0; e < 18; e++) acc[e] += term[e];
  }
  for (int i = 0; i < s; i++) mmul(acc, acc, acc);
  memcpy(out, acc, sizeof acc);
}

// Gram-Schmidt on rows, then det phase fix on the last column.
void reunit(double *m) {
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < i; j++) {
      // z = <row j, row i> = sum conj(m_j c) m_i c ; row i -= z row j
      double zr = 0, zi = 0;
      for (int c = 0; c < 3; c++) {
        double ar = m[6 * j + 2 * c], ai = m[6 * j + 2 * c + 1];
        double br = m[6 * i + 2 * c], bi = m[6 * i + 2 * c + 1];
        zr += ar * br + ai * bi;
        zi += ar * bi - ai * br;
      }
      for (int c = 0; c < 3; c++) {
        double ar = m[6 * j + 2 * c], ai = m[6 * j + 2 * c + 1];
        m[6 * i + 2 * c] -= zr * ar - zi * ai;
        m[6 * i + 2 * c + 1] -= zr * ai + zi * ar;
      }
    }
    double nn = 0;
    for (int c = 0; c < 6; c++) nn += m[6 * i + c] * m[6 * i + c];
    nn = 1.0 / std::sqrt(nn);
    for (int c = 0; c < 6; c++) m[6 * i + c] *= nn;
  }
  // determinant by cofactor expansion
  auto C = [&](int i, int j, double &re, double &im) { re = m[6 * i + 2 * j]; im = m[6 * i + 2 * j + 1]; };
  double dr = 0, di = 0;
  for (int j = 0; j < 3; j++) {
    int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    double ar, ai, br, bi, cr, ci, er, ei;
    C(1, j1, ar, ai); C(2, j2, br, bi); C(1, j2, cr, ci); C(2, j1, er, ei);
    double mr = ar * br - ai * bi - (cr * er - ci * ei);
    double mi = ar * bi + ai * br - (cr * ei + ci * er);
    double xr, xi;
    C(0, j, xr, xi);
    dr += xr * mr - xi * mi;
    di += xr * mi + xi * mr;
  }
  double r = std::sqrt(dr * dr + di * di);
  double pr = dr / r, pi = -di / r;  // conj(det)/|det|
  for (int c = 0; c < 3; c++) {
    double a = m[6 * c + 4], b = m[6 * c + 5];
    m[6 * c + 4] = a * pr - b * pi;
    m[6 * c + 5] = a * pi + b * pr;
  }
}

template <typename F>
void parallel_for(long n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 32) nt = 32;
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([=]() {
      long b = n * t / nt, e = n * (t + 1) / nt;
      for (long i = b; i < e; i++) f(i);
    });
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

void QOPX_gen_gauge(const int L[4], unsigned long long seed, double eps, double *U[4]) {
  long V = (long)L[0] * L[1] * L[2] * L[3];
  parallel_for(V, [&](long i) {
    for (int mu = 0; mu < 4; mu++) {
      SplitMix rng(seed, (uint64_t)i * 4 + mu);
      double G[18];
      for (int e = 0; e < 18; e += 2) rng.gauss2(G[e], G[e + 1]);
      // A = (G + G^+)/2 ; H = A - tr(A)/3 ; X = i eps H
      double H[18];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          H[6 * a + 2 * b] = 0.5 * (G[6 * a + 2 * b] + G[6 * b + 2 * a]);
          H[6 * a + 2 * b + 1] = 0.5 * (G[6 * a + 2 * b + 1] - G[6 * b + 2 * a + 1]);
        }
      double tr = (H[0] + H[8] + H[16]) / 3.0;
      for (int a = 0; a < 3; a++) H[8 * a] -= tr;
      double X[18];
      for (int e = 0; e < 18; e += 2) { X[e] = -eps * H[e + 1]; X[e + 1] = eps * H[e]; }
      double *m = U[mu] + 18 * i;
      mexp(m, X);
      reunit(m);
    }
  });
}

void QOPX_gen_gaussian_V(const int L[4], unsigned long long seed, int evenodd, double *v) {
  long V = (long)L[0] * L[1] * L[2] * L[3], Vh = V / 2;
  const double s = std::sqrt(0.5);  // unit variance per complex component
  parallel_for(V, [&](long i) {
    bool on = evenodd == 2 || (evenodd == 0 && i < Vh) || (evenodd == 1 && i >= Vh);
    if (!on) {
      for (int c = 0; c < 6; c++) v[6 * i + c] = 0;
      return;
    }
    SplitMix rng(seed, 0x100000000ull + (uint64_t)i);
    for (int c = 0; c < 6; c += 2) {
      rng.gauss2(v[6 * i + c], v[6 * i + c + 1]);
      v[6 * i + c] *= s;
      v[6 * i + c + 1] *= s;
    }
  });
}

}  // extern "C"
