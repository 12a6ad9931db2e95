// eigCG (lib/invert_eigcg_p.c:1-1068; Stathopoulos and Orginos,
// arXiv:0707.0131) on the device: the CG iterates, the Lanczos vectors and
// the deflation space live in HBM; the small dense eigenproblems (the
// Lanczos matrix T, at most eigcg_m square, and the Rayleigh-Ritz matrix,
// at most eigcg_l square) are solved on the host by cyclic Jacobi, where the
// reference calls LAPACK's dsyev / zheev; eigenvectors are defined up to
// sign or phase, which the algorithm does not depend on.
#include "eigcg.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <utility>

#include "ops.h"

namespace qop {

using zd = std::complex<double>;

EigcgState::~EigcgState() {
  if (store) {
    sync();
    QOP_CHECK_HIP(hipFree(store));
  }
}

namespace {

// symmetric a (n x n row-major, destroyed) -> w ascending, z columns
void jacobi_sym(int n, std::vector<double> &a, std::vector<double> &w, std::vector<double> &z) {
  z.assign((size_t)n * n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; i++) z[(size_t)i * n + i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double v = a[(size_t)i * n + j] * a[(size_t)i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off == 0 || off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0) continue;
        const double th = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          double &kp = a[(size_t)k * n + p], &kq = a[(size_t)k * n + q];
          const double x = kp, y = kq;
          kp = c * x - s * y;
          kq = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          double &pk = a[(size_t)p * n + k], &qk = a[(size_t)q * n + k];
          const double x = pk, y = qk;
          pk = c * x - s * y;
          qk = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          double &kp = z[(size_t)k * n + p], &kq = z[(size_t)k * n + q];
          const double x = kp, y = kq;
          kp = c * x - s * y;
          kq = s * x + c * y;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = a[(size_t)i * n + i];
  for (int i = 0; i < n; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (w[j] < w[k]) k = j;
    if (k != i) {
      std::swap(w[i], w[k]);
      for (int r = 0; r < n; r++) std::swap(z[(size_t)r * n + i], z[(size_t)r * n + k]);
    }
  }
}

// Hermitian a (destroyed) -> w ascending, z columns
void jacobi_herm(int n, std::vector<zd> &a, std::vector<double> &w, std::vector<zd> &z) {
  z.assign((size_t)n * n, zd(0));
  w.assign(n, 0.0);
  for (int i = 0; i < n; i++) z[(size_t)i * n + i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double v = std::norm(a[(size_t)i * n + j]);
        tot += v;
        if (i != j) off += v;
      }
    if (off == 0 || off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const zd apq = a[(size_t)p * n + q];
        const double g = std::abs(apq);
        if (g == 0) continue;
        const zd ph = apq / g;
        const double th = (a[(size_t)q * n + q].real() - a[(size_t)p * n + p].real()) / (2 * g);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        // columns p' = c e_p - s conj(ph) e_q, q' = s ph e_p + c e_q
        for (int k = 0; k < n; k++) {
          zd &kp = a[(size_t)k * n + p], &kq = a[(size_t)k * n + q];
          const zd x = kp, y = kq;
          kp = c * x - s * std::conj(ph) * y;
          kq = s * ph * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          zd &pk = a[(size_t)p * n + k], &qk = a[(size_t)q * n + k];
          const zd x = pk, y = qk;
          pk = c * x - s * ph * y;
          qk = s * std::conj(ph) * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          zd &kp = z[(size_t)k * n + p], &kq = z[(size_t)k * n + q];
          const zd x = kp, y = kq;
          kp = c * x - s * std::conj(ph) * y;
          kq = s * ph * x + c * y;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = a[(size_t)i * n + i].real();
  for (int i = 0; i < n; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (w[j] < w[k]) k = j;
    if (k != i) {
      std::swap(w[i], w[k]);
      for (int r = 0; r < n; r++) std::swap(z[(size_t)r * n + i], z[(size_t)r * n + k]);
    }
  }
}

class Eig {
 public:
  Eig(Links &L, double mass, int eo, EigcgState &eg)
      : L_(L), mass_(mass), eo_(eo), eg_(eg), Vh_(L.lat->Vh), tmp_(L.lat, P64) {}

  // parity-half slot i of the store as a full-vector view on subset eo
  DevVec view(int i) const {
    DevVec v;
    v.lat = L_.lat;
    v.prec = P64;
    // the store leads with one spare half, so the view's other-parity half
    // (never touched: every operation here is on subset eo) stays inside it
    v.data = (char *)slot(i) - (eo_ ? (size_t)3 * Vh_ * sizeof(double2) : 0);
    return v;
  }
  double linop(DevVec &o, const DevVec &i) { return ddagd_norm2(L_, P64, mass_, o, i, eo_, tmp_.v); }

  // y_j = sum_i R(i, j) x_i for j < n2 (R host, row-major n1 x n2) with
  // the pointer arrays and R staged in one device buffer kept for the solve
  void combine(const double2 *const *xp, int n1, double2 *const *yp, int n2, const std::vector<zd> &R) {
    const size_t nb = (n1 + n2) * sizeof(void *) + (size_t)n1 * n2 * sizeof(double2);
    if (nb > args_cap_) {
      sync();
      DevBuf<char> grown(nb);  // the old buffer is freed when `grown` goes
      std::swap(args_.p, grown.p);
      args_cap_ = nb;
    }
    std::vector<char> h(nb);
    memcpy(h.data(), xp, n1 * sizeof(void *));
    memcpy(h.data() + n1 * sizeof(void *), yp, n2 * sizeof(void *));
    memcpy(h.data() + (n1 + n2) * sizeof(void *), R.data(), (size_t)n1 * n2 * sizeof(double2));
    // synchronous upload: the buffer is reused by the next call
    QOP_CHECK_HIP(hipStreamSynchronize(stream()));
    QOP_CHECK_HIP(hipMemcpy(args_.p, h.data(), nb, hipMemcpyHostToDevice));
    launch_rotate((const double2 *const *)args_.p, (double2 *const *)(args_.p + n1 * sizeof(void *)),
                  (const double2 *)(args_.p + (n1 + n2) * sizeof(void *)), n1, n2, n2, 3 * Vh_, stream());
  }

  // slots s[j] = sum_i R(i, j) slots s[i] for j < n2 (R host, row-major n1 x n2)
  void rotate(const int *s, int n1, int n2, const std::vector<zd> &R) {
    std::vector<const double2 *> xp(n1);
    std::vector<double2 *> yp(n2);
    for (int i = 0; i < n1; i++) xp[i] = slot(s[i]);
    for (int j = 0; j < n2; j++) yp[j] = slot(eg_.cap + j);
    combine(xp.data(), n1, yp.data(), n2, R);
    for (int j = 0; j < n2; j++)
      QOP_CHECK_HIP(hipMemcpyAsync(slot(s[j]), yp[j], 3 * Vh_ * sizeof(double2),
                                   hipMemcpyDeviceToDevice, stream()));
    sync();
  }

  double2 *slot(int i) const { return eg_.store + (size_t)(i + 1) * 3 * Vh_; }

  // G(i, j) = <slot x0+i, slot y0+j> (n1 x n2, host row-major), in tiles
  // the Gram kernel takes
  std::vector<zd> gram(int x0, int n1, int y0, int n2) {
    constexpr int NBLK = 256;
    if (!partial_.p) dev_alloc(&partial_.p, (size_t)NBLK * gram_max_pairs() * sizeof(double2));
    if (!gdev_.p) dev_alloc(&gdev_.p, (size_t)gram_max_pairs() * sizeof(double2));
    std::vector<zd> G((size_t)n1 * n2);
    const int cj0 = std::min(n2, gram_max_rows() / 2);
    for (int j0 = 0; j0 < n2; j0 += cj0) {
      const int cj = std::min(cj0, n2 - j0);
      const int ci0 = std::max(1, std::min(gram_max_rows() - cj, gram_max_pairs() / cj));
      for (int i0 = 0; i0 < n1; i0 += ci0) {
        const int ci = std::min(ci0, n1 - i0);
        launch_gram(slot(x0 + i0), slot(y0 + j0), 3 * Vh_, ci, cj, 3 * Vh_, partial_.p, NBLK, gdev_.p, stream());
        std::vector<double2> h((size_t)ci * cj);
        QOP_CHECK_HIP(hipMemcpyAsync(h.data(), gdev_.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost,
                                     stream()));
        sync();
        for (int i = 0; i < ci; i++)
          for (int j = 0; j < cj; j++)
            G[(size_t)(i0 + i) * n2 + j0 + j] = zd(h[(size_t)i * cj + j].x, h[(size_t)i * cj + j].y);
      }
    }
    return G;
  }

  // slots s0+i0 .. s0+i0+n-1 orthonormalised against all slots before them
  // (invert_eigcg_p.c:201-213) in block form: the new block is projected off
  // the i0 earlier slots twice (one Gram product and one combination per
  // pass), then orthonormalised within itself by Cholesky QR, twice; the
  // reference subtracts vector by vector.  The span is the same; a block
  // whose Gram matrix is not numerically positive definite falls back to
  // the vector-by-vector form.
  void orthonormalize(int s0, int i0, int n) {
    if (n <= 0) return;
    for (int pass = 0; pass < 2 && i0 > 0; pass++) {
      const std::vector<zd> c = gram(s0, i0, s0 + i0, n);  // i0 x n
      std::vector<int> sl(n + i0);
      std::vector<zd> R((size_t)(n + i0) * n, zd(0));
      for (int j = 0; j < n; j++) {
        sl[j] = s0 + i0 + j;
        R[(size_t)j * n + j] = 1;
      }
      for (int k = 0; k < i0; k++) {
        sl[n + k] = s0 + k;
        for (int j = 0; j < n; j++) R[(size_t)(n + k) * n + j] = -c[(size_t)k * n + j];
      }
      rotate(sl.data(), n + i0, n, R);
    }
    bool ok = true;
    for (int pass = 0; pass < 2 && ok; pass++) {
      std::vector<zd> G = gram(s0 + i0, n, s0 + i0, n);  // Hermitian, G(i, j) = <v_i, v_j>
      // G = U^+ U (U upper triangular); the block times U^-1 is orthonormal
      std::vector<zd> U((size_t)n * n, zd(0));
      for (int j = 0; j < n && ok; j++) {
        zd d = G[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= std::conj(U[(size_t)k * n + j]) * U[(size_t)k * n + j];
        if (!(d.real() > 1e-12 * std::abs(G[(size_t)j * n + j]))) { ok = false; break; }
        const double ujj = std::sqrt(d.real());
        U[(size_t)j * n + j] = ujj;
        for (int i = j + 1; i < n; i++) {
          zd e = G[(size_t)j * n + i];
          for (int k = 0; k < j; k++) e -= std::conj(U[(size_t)k * n + j]) * U[(size_t)k * n + i];
          U[(size_t)j * n + i] = e / ujj;
        }
      }
      if (!ok) break;
      // W = U^-1 (upper triangular), by back substitution column by column
      std::vector<zd> W((size_t)n * n, zd(0));
      for (int j = 0; j < n; j++) {
        W[(size_t)j * n + j] = 1.0 / U[(size_t)j * n + j];
        for (int i = j - 1; i >= 0; i--) {
          zd sum = 0;
          for (int k = i + 1; k <= j; k++) sum += U[(size_t)i * n + k] * W[(size_t)k * n + j];
          W[(size_t)i * n + j] = -sum / U[(size_t)i * n + i];
        }
      }
      std::vector<int> sl(n);
      for (int j = 0; j < n; j++) sl[j] = s0 + i0 + j;
      rotate(sl.data(), n, n, W);  // v'_j = sum_i W(i, j) v_i
    }
    if (ok) return;
    for (int i = i0; i < i0 + n; i++) {
      for (int pass = 0; pass < 2 && i > 0; pass++) {
        const std::vector<zd> c = gram(s0, i, s0 + i, 1);
        std::vector<int> sl(i + 1);
        std::vector<zd> R(i + 1);
        sl[0] = s0 + i;
        R[0] = 1;
        for (int k = 0; k < i; k++) {
          sl[k + 1] = s0 + k;
          R[k + 1] = -c[k];
        }
        rotate(sl.data(), i + 1, 1, R);
      }
      DevVec vi = view(s0 + i);
      const double r = 1.0 / std::sqrt(norm2(vi, eo_));
      scal(vi, Cplx{r, 0}, vi, eo_);
    }
  }

  // invert_eigcg_p.c:215-326 (without the verbose residual report)
  void rayleigh_ritz(int nu, int nv, DevVec &Mp) {
    const int n = nu + nv;
    if (n <= 0) return;
    (void)Mp;
    std::vector<int> s(n);
    for (int i = 0; i < n; i++) s[i] = i;
    orthonormalize(0, nu, nv);
    std::vector<zd> mm((size_t)n * n, zd(0)), tt;
    std::vector<double> vv;
    for (int i = 0; i < nu; i++) mm[(size_t)i * n + i] = eg_.l[i];
    // A u_j of the new vectors into the scratch slots, then one Gram matrix
    for (int j = nu; j < n; j++) {
      DevVec uj = view(j), w = view(eg_.cap + j - nu);
      linop(w, uj);
    }
    const std::vector<zd> G = gram(0, n, eg_.cap, nv);
    for (int j = nu; j < n; j++)
      for (int i = 0; i <= j; i++) {
        const zd z = G[(size_t)i * nv + (j - nu)];
        mm[(size_t)i * n + j] = z;
        if (i < j) mm[(size_t)j * n + i] = std::conj(z);
      }
    jacobi_herm(n, mm, vv, tt);
    rotate(s.data(), n, n, tt);
    for (int i = 0; i < n; i++) eg_.l[i] = vv[i];
    if (n == 1) eg_.l[1] = 0;
  }

  // x = sum_i u_i (u_i^+ b) / l_i (invert_eigcg_p.c:328-341): the n
  // products in one Gram pass (b staged in a scratch slot), the sum in one
  // combination
  void initcg(DevVec &x, int n, const DevVec &b) {
    if (n <= 0) return;
    QOP_CHECK_HIP(hipMemcpyAsync(slot(eg_.cap), b.half(eo_).data, 3 * Vh_ * sizeof(double2),
                                 hipMemcpyDeviceToDevice, stream()));
    const std::vector<zd> z = gram(0, n, eg_.cap, 1);
    std::vector<zd> c(n);
    for (int i = 0; i < n; i++) c[i] = z[i] / eg_.l[i];
    std::vector<const double2 *> xp(n);
    for (int i = 0; i < n; i++) xp[i] = slot(i);
    double2 *yp[1] = {(double2 *)x.half(eo_).data};
    combine(xp.data(), n, yp, 1, c);
  }

  int nev, m;
  std::vector<double> t, tb, td, y;  // m x m, m x m, m, m x 2 nev

  // td[0:nv], tb[0:nv,0:nv] = eigenpairs of t[0:nv,0:nv] (:488-501)
  void diag_t(int nv) {
    std::vector<double> a((size_t)nv * nv), w, z;
    for (int i = 0; i < nv; i++)
      for (int j = 0; j < nv; j++) a[(size_t)i * nv + j] = t[(size_t)i * m + j];
    jacobi_sym(nv, a, w, z);
    for (int i = 0; i < nv; i++) td[i] = w[i];
    for (int i = 0; i < nv; i++)
      for (int j = 0; j < nv; j++) tb[(size_t)i * m + j] = z[(size_t)i * nv + j];
  }
  // invert_eigcg_p.c:503-550
  void set_y(int noff, int ne, int nv, int n) {
    const int ly = 2 * nev;
    for (int i = 0; i < nv; i++)
      for (int j = 0; j < ne; j++) y[(size_t)i * ly + noff + j] = tb[(size_t)i * m + j];
    for (int i = nv; i < n; i++)
      for (int j = 0; j < ne; j++) y[(size_t)i * ly + noff + j] = 0;
    if (!noff) return;
    for (int c = noff; c < noff + ne; c++) {
      for (int k = 0; k < c; k++) {
        double d = 0;
        for (int i = 0; i < n; i++) d += y[(size_t)i * ly + k] * y[(size_t)i * ly + c];
        for (int i = 0; i < n; i++) y[(size_t)i * ly + c] -= d * y[(size_t)i * ly + k];
      }
      double q = 0;
      for (int i = 0; i < n; i++) q += y[(size_t)i * ly + c] * y[(size_t)i * ly + c];
      q = 1 / std::sqrt(q);
      for (int i = 0; i < n; i++) y[(size_t)i * ly + c] *= q;
    }
    const int k = noff + ne;
    std::vector<double> tb1((size_t)n * k), h0((size_t)k * k), w, ta;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < k; j++) {
        double s2 = 0;
        for (int q = 0; q < n; q++) s2 += t[(size_t)i * m + q] * y[(size_t)q * ly + j];
        tb1[(size_t)i * k + j] = s2;
      }
    for (int i = 0; i < k; i++)
      for (int j = 0; j < k; j++) {
        double s2 = 0;
        for (int q = 0; q < n; q++) s2 += y[(size_t)q * ly + i] * tb1[(size_t)q * k + j];
        h0[(size_t)i * k + j] = s2;
      }
    jacobi_sym(k, h0, w, ta);
    for (int i = 0; i < k; i++) td[i] = w[i];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < k; j++) {
        double s2 = 0;
        for (int q = 0; q < k; q++) s2 += y[(size_t)i * ly + q] * ta[(size_t)q * k + j];
        tb[(size_t)i * m + j] = s2;
      }
  }
  // v[0:n] = v[0:nv] tb[0:nv, 0:n] (:552-559); v = slots v0 ..
  void rot_t(int v0, int n, int nv) {
    if (n > nv) n = nv;
    std::vector<zd> R((size_t)nv * n);
    for (int i = 0; i < nv; i++)
      for (int j = 0; j < n; j++) R[(size_t)i * n + j] = tb[(size_t)i * m + j];
    std::vector<int> s(nv);
    for (int i = 0; i < nv; i++) s[i] = v0 + i;
    rotate(s.data(), nv, n, R);
  }
  void setv(int v0, int &nv, const DevVec &r, double rsq) {
    DevVec d = view(v0 + nv);
    scal(d, Cplx{1 / std::sqrt(rsq), 0}, r, eo_);
    nv++;
  }
  // t(i, j) = Re <v_i, A v_j> (:363-394): A v_j into the scratch slots, then
  // one Gram matrix
  void gett(int v0, int n, DevVec &Mp) {
    (void)Mp;
    for (int j = 0; j < n; j++) {
      DevVec vj = view(v0 + j), w = view(eg_.cap + j);
      linop(w, vj);
    }
    const std::vector<zd> G = gram(v0, n, eg_.cap, n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) t[(size_t)i * m + j] = G[(size_t)i * n + j].real();
  }
  // invert_eigcg_p.c:561-616
  void reset_t(int v0, int &nv, double a0, double a, double b, const DevVec &r, double rsq, DevVec &Mp) {
    nv = 2 * nev;
    setv(v0, nv, r, rsq);
    std::fill(t.begin(), t.end(), 0.0);
    DevVec vl = view(v0 + nv - 1);
    linop(Mp, vl);
    // Re <v_i, A v_last> for i < nv - 1 in one Gram pass (A v_last staged in
    // a scratch slot)
    QOP_CHECK_HIP(hipMemcpyAsync(slot(eg_.cap), Mp.half(eo_).data, 3 * Vh_ * sizeof(double2),
                                 hipMemcpyDeviceToDevice, stream()));
    const std::vector<zd> g = gram(v0, nv - 1, eg_.cap, 1);
    for (int i = 0; i < nv - 1; i++) {
      const double rr = g[i].real();
      t[(size_t)i * m + nv - 1] = rr;
      t[(size_t)(nv - 1) * m + i] = rr;
      t[(size_t)i * m + i] = td[i];
    }
    t[(size_t)(nv - 1) * m + nv - 1] = 1. / a + b / a0;
  }

 private:
  Links &L_;
  double mass_;
  int eo_;
  EigcgState &eg_;
  long Vh_;
  TmpVec tmp_;
  DevBuf<double2> partial_, gdev_;
  DevBuf<char> args_;  // combine(): pointer arrays and coefficients
  size_t args_cap_ = 0;
};

}  // namespace

void invert_eigcg(Links &L, double mass, int eo, InvArg *ia, ResArg *ra, DevVec &out,
                  const DevVec &in, EigcgState &eg) {
  Lattice *lat = L.lat;
  double a = 0, b = 0, a0 = 0, rsq, relnorm = 1.0, oldrsq, pkp, insq, rsqstop;
  int iteration = 0, total = 0, nrestart = -1;
  const int max_iterations = ia->max_iter;
  int max_restarts = ia->max_restarts;
  if (max_restarts < 0) max_restarts = 5;
  const int nev = eg.nev, m = eg.m, numax = eg.numax;
  if (nev < 1 || m < 2 * nev + 2 || numax < m)
    throw std::runtime_error("eigCG: needs eigcg_nev >= 1, eigcg_m >= 2 nev + 2, eigcg_l >= eigcg_m");
  if (!eg.store) {
    eg.cap = numax;
    eg.lat = lat;
    dev_alloc(&eg.store, (size_t)(2 * numax + 1) * 3 * lat->Vh * sizeof(double2));
    eg.l.assign(numax + 1, 0.0);
    eg.nu = eg.nv = eg.nn = 0;
    eg.addvecs = 1;
  }
  if (eg.eo != eo) {
    // the vectors live on the parity of the solves that made them; a solve
    // on the other parity starts a new space
    eg.nu = eg.nv = eg.nn = 0;
    eg.addvecs = 1;
    eg.eo = eo;
  }
  int nu = eg.nu, nv = eg.nv, nn = eg.nn, addvecs = eg.addvecs;
  // the reference writes the Lanczos vectors to u[nu + nn ..] without a
  // bound; stop collecting when they would not fit
  if (addvecs && nu + nn + m > eg.cap) addvecs = 0;
  Eig E(L, mass, eo, eg);
  E.nev = nev;
  E.m = m;
  E.t.assign((size_t)m * m, 0.0);
  E.tb.assign((size_t)m * m, 0.0);
  E.td.assign(m + 1, 0.0);
  E.y.assign((size_t)m * 2 * nev, 0.0);
  TmpVec r(lat, P64), Mp(lat, P64), Mq(lat, P64), p(lat, P64);
  int v0 = 0;
  insq = norm2(in, eo);
  rsqstop = ra->rsqmin * insq;
  rsq = 0;
  ra->final_rsq = 0; ra->final_rel = 0; ra->final_iter = 0; ra->final_restart = 0;
  if (insq == 0.) {
    zero(out, eo);
    return;
  }
  auto done = [&] {
    return ((rsqstop <= 0 || rsq < rsqstop) && (ra->relmin <= 0 || relnorm < ra->relmin)) ||
           total >= max_iterations;
  };
  do {
    copy(p.v, out, eo);
    E.linop(Mp.v, p.v);
    iteration = 1;
    total++;
    axpy(r.v, -1.0, Mp.v, in, eo);
    rsq = norm2(r.v, eo);
    if (done()) break;
    if (addvecs && nn >= m) {
      int nr = numax - 2 * m - 1;
      if (nr >= nu) nr = nu - 2;
      if (nr < 0) nr = 0;
      const double oldl = nu == 0 ? 9999 : eg.l[nr];
      E.rayleigh_ritz(nu, nn, Mq.v);
      nu += nn;
      nn = 0;
      const double delta = std::fabs(2 * (oldl - eg.l[nr]) / (oldl + eg.l[nr]));
      if (delta < 1e-4) addvecs = 0;
      else if (nu > numax - 2 * m) nu = numax - 2 * m;
    }
    if (nu > 0) {
      E.initcg(p.v, nu, r.v);
      axpy(out, 1.0, p.v, out, eo);
      copy(p.v, out, eo);
      E.linop(Mp.v, p.v);
      iteration = 1;
      total++;
      axpy(r.v, -1.0, Mp.v, in, eo);
      rsq = norm2(r.v, eo);
      if (ra->relmin > 0) relnorm = relnorm2(r.v, out, eo);
      if (done()) break;
    }
    if (addvecs) {
      v0 = nu + nn;
      std::fill(E.t.begin(), E.t.end(), 0.0);
    }
    copy(p.v, r.v, eo);
    while (true) {
      E.linop(Mp.v, p.v);
      iteration++;
      total++;
      pkp = redot(p.v, Mp.v, eo);
      a0 = a;
      a = rsq / pkp;
      if (addvecs) {
        if (nv == 0) {
          E.t[0] = 1. / a;
          E.setv(v0, nv, r.v, rsq);
        } else if (nv < m) {
          const double td = 1. / a + b / a0, ts = -std::sqrt(b) / a0;
          E.t[(size_t)nv * m + nv] = td;
          E.t[(size_t)nv * m + nv - 1] = ts;
          E.t[(size_t)(nv - 1) * m + nv] = ts;
          E.setv(v0, nv, r.v, rsq);
        } else {
          E.diag_t(nv);
          E.set_y(0, nev, nv, nv);
          E.diag_t(nv - 1);
          E.set_y(nev, nev, nv - 1, nv);
          E.rot_t(v0, 2 * nev, nv);
          E.reset_t(v0, nv, a0, a, b, r.v, rsq, Mq.v);
        }
      }
      axpy(out, a, p.v, out, eo);
      axpy(r.v, -a, Mp.v, r.v, eo);
      oldrsq = rsq;
      rsq = norm2(r.v, eo);
      if (ra->relmin > 0) relnorm = relnorm2(r.v, out, eo);
      if (done()) break;
      b = rsq / oldrsq;
      axpy(p.v, b, p.v, r.v, eo);
    }
    if (addvecs) {
      if (nv > nev) {
        E.orthonormalize(v0, 0, nv);
        E.gett(v0, nv, Mq.v);
        E.diag_t(nv);
        E.set_y(0, nev, nv, nv);
        E.rot_t(v0, nev, nv);
        nn += nev;
      }
      nv = 0;
    }
    nrestart++;
  } while (total < max_iterations && nrestart < max_restarts);
  if (addvecs) {
    E.rayleigh_ritz(0, nu + nn + nv, Mq.v);
    nu += nn + nv;
    nn = 0;
    nv = 0;
  }
  eg.nu = nu;
  eg.nv = nv;
  eg.nn = nn;
  eg.addvecs = addvecs;
  ra->final_rsq = rsq / insq;
  ra->final_rel = relnorm;
  ra->final_iter = total;
  ra->final_restart = nrestart;
}

}  // namespace qop
