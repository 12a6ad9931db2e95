// Runtime pieces of the MI355X asqtad library: error handling, lattice
// cache, device allocation, the library stream and reduction scratch.
#include <map>
#include <mutex>

#include "common.h"
#include "eigcg.h"

namespace qop {

[[noreturn]] void fatal(const char *file, int line, const char *msg) {
  fprintf(stderr, "QOP-MI355X fatal error at %s:%d: %s\n", file, line, msg);
  fflush(stderr);
  abort();
}

void dev_alloc_bytes(void **p, size_t bytes) {
  *p = nullptr;
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();  // clear the error state of this thread
    *p = nullptr;
    throw DeviceOutOfMemory{bytes};
  }
  if (e != hipSuccess) fatal(__FILE__, __LINE__, hipGetErrorString(e));
}

static std::mutex g_mu;

Lattice *lattice_get(const int L[4]) {
  static std::map<std::vector<int>, Lattice *> cache;
  std::lock_guard<std::mutex> lk(g_mu);
  std::vector<int> key(L, L + 4);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  for (int i = 0; i < 4; i++)
    QOP_REQUIRE(L[i] >= 4 && L[i] % 2 == 0, "lattice extents must be even and >= 4");
  Lattice *lat = new Lattice;
  for (int i = 0; i < 4; i++) lat->L[i] = L[i];
  lat->V = (long)L[0] * L[1] * L[2] * L[3];
  QOP_REQUIRE(lat->V < (1L << 31), "lattice volume must fit in a 32-bit index");
  lat->Vh = lat->V / 2;
  Geom &g = lat->g;
  for (int i = 0; i < 4; i++) g.L[i] = L[i];
  g.Lh3 = L[3] / 2;
  g.Vh = (int)lat->Vh;
  g.stride[3] = 1;
  g.stride[2] = L[3];
  g.stride[1] = L[3] * L[2];
  g.stride[0] = L[3] * L[2] * L[1];
  cache[key] = lat;
  return lat;
}

void devvec_alloc(DevVec &v, Lattice *lat, int prec) {
  v.lat = lat;
  v.prec = prec;
  v.norm = nullptr;
  dev_alloc(&v.data, (size_t)2 * 3 * lat->Vh * prec_cbytes(prec));
  QOP_CHECK_HIP(hipMemsetAsync(v.data, 0, (size_t)2 * 3 * lat->Vh * prec_cbytes(prec), stream()));
  if (prec == P16) {
    try {
      dev_alloc(&v.norm, (size_t)2 * lat->Vh * sizeof(float));
    } catch (const DeviceOutOfMemory &) {
      devvec_free(v);
      throw;
    }
    QOP_CHECK_HIP(hipMemsetAsync(v.norm, 0, (size_t)2 * lat->Vh * sizeof(float), stream()));
  }
}

void devvec_free(DevVec &v) {
  if (v.data) QOP_CHECK_HIP(hipFree(v.data));
  if (v.norm) QOP_CHECK_HIP(hipFree(v.norm));
  v.data = nullptr;
  v.norm = nullptr;
}

void devbatch_alloc(DevBatch &v, Lattice *lat, int prec, int n) {
  v.lat = lat;
  v.prec = prec;
  v.n = n;
  size_t bytes = (size_t)n * 3 * lat->Vh * prec_cbytes(prec);
  v.norm = nullptr;
  dev_alloc(&v.data, bytes);
  QOP_CHECK_HIP(hipMemsetAsync(v.data, 0, bytes, stream()));
  if (prec == P16) {
    try {
      dev_alloc(&v.norm, (size_t)n * lat->Vh * sizeof(float));
    } catch (const DeviceOutOfMemory &) {
      devbatch_free(v);
      throw;
    }
    QOP_CHECK_HIP(hipMemsetAsync(v.norm, 0, (size_t)n * lat->Vh * sizeof(float), stream()));
  }
}

void devbatch_free(DevBatch &v) {
  if (v.data) QOP_CHECK_HIP(hipFree(v.data));
  if (v.norm) QOP_CHECK_HIP(hipFree(v.norm));
  v.data = nullptr;
  v.norm = nullptr;
  v.n = 0;
}

void linkset_free(LinkSet &l) {
  if (l.data) QOP_CHECK_HIP(hipFree(l.data));
  if (l.scale) QOP_CHECK_HIP(hipFree(l.scale));
  l.data = nullptr;
  l.scale = nullptr;
  l.prec = -1;
}

LinkSet &Links::get(int prec) {
  if (prec == P64) return l64;
  if (prec == P32H) prec = P16;  // FP32 vectors read the FP16 link set
  LinkSet &d = prec == P32 ? l32 : l16;
  if (!d.data) {
    convert_links(lat, l64, d, prec, stream());
    sync();  // the converted set may be read from any stream
  }
  return d;
}

LinkSet &Links::get_single(int prec) {
  QOP_REQUIRE(prec == P64 || prec == P32, "single-stored links exist in FP64 and FP32");
  LinkSet &d = prec == P64 ? s64 : s32;
  if (!d.data) {
    // copy the forward blocks (p, k = 4mu+{0,1} or 2mu) of the double-stored set
    LinkSet &src = get(prec);
    const int nls = nl / 2, per = nl / 4;
    const size_t blk = (size_t)9 * lat->Vh * prec_cbytes(prec);
    dev_alloc(&d.data, (size_t)2 * nls * blk);
    d.prec = prec;
    d.nl = nls;
    for (int p = 0; p < 2; p++)
      for (int mu = 0; mu < 4; mu++)
        for (int j = 0; j < per / 2; j++) {
          const int ks = (per / 2) * mu + j, kd = per * mu + j;
          QOP_CHECK_HIP(hipMemcpyAsync((char *)d.data + ((size_t)p * nls + ks) * blk,
                                       (const char *)src.data + ((size_t)p * nl + kd) * blk, blk,
                                       hipMemcpyDeviceToDevice, stream()));
        }
    sync();
  }
  return d;
}

void Links::drop_derived() {
  linkset_free(l32);
  linkset_free(l16);
  linkset_free(s64);
  linkset_free(s32);
}

void Links::free_all() {
  linkset_free(l64);
  drop_derived();
  delete eig;
  eig = nullptr;
}

hipStream_t stream() {
  static hipStream_t s = [] {
    hipStream_t t;
    QOP_CHECK_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    return t;
  }();
  return s;
}
void sync() { QOP_CHECK_HIP(hipStreamSynchronize(stream())); }

namespace {
struct Scratch {
  void *p = nullptr;
  size_t n = 0;
  bool host = false;
  void *get(size_t bytes) {
    if (bytes <= n) return p;
    if (p) {
      QOP_CHECK_HIP(hipStreamSynchronize(stream()));
      if (host) QOP_CHECK_HIP(hipHostFree(p)); else QOP_CHECK_HIP(hipFree(p));
    }
    p = nullptr;
    n = 0;
    const size_t want = bytes < (1u << 20) ? (1u << 20) : bytes * 2;
    if (host) QOP_CHECK_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
    else dev_alloc(&p, want);
    n = want;
    return p;
  }
};
Scratch s_partial, s_result, s_host{nullptr, 0, true};
}  // namespace

double *red_partial(size_t nd) { return (double *)s_partial.get(nd * sizeof(double)); }
double *red_result(size_t nd) { return (double *)s_result.get(nd * sizeof(double)); }
double *red_host(size_t nd) { return (double *)s_host.get(nd * sizeof(double)); }

}  // namespace qop
