// Allocation-free multi-RHS hop (QOPX_D3_asqtad_hop_noalloc): the same
// conversion, hop and write-back kernels as QOPX_D3_asqtad_hop, with the
// stored-precision source and result batches placed in a fixed device array
// of this file instead of device memory allocated per call.  For callers that
// cannot allocate (code under stream capture, serving loops that must not
// reach the allocator).
//
// The array lives for the whole process on every device the library runs on.
// Its device address is looked up once per device (hipGetSymbolAddress, on
// the entry point's first call with that device current) and only that
// address is passed to kernels and copies; the array's name in host code is
// a host-side stand-in, not a device address.  Calls on one device share the
// array: calls that overlap in time on one device race, and nothing of one
// call's contents is meant to survive into the next.
#include <mutex>

#include "batch.h"
#include "handles.h"

namespace qop {

// 64 MiB: e.g. 16 FP64 vectors in and out on a 16^4 lattice (2 x 16 x 48 B x
// 32768 sites of a parity = 48 MiB), or 64 FP16 vectors on 24^3x48 halves.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
void *scratch_ptr() {
  int dev = 0;
  QOP_CHECK_HIP(hipGetDevice(&dev));
  QOP_REQUIRE(dev >= 0 && dev < kMaxDevices, "hop_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw std::runtime_error(std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// bytes of one stored batch of n vectors on one parity (data, then FP16 norms)
size_t batch_bytes(const Lattice *lat, int p, int n, size_t *norm_off) {
  const size_t data = round256((size_t)n * 3 * lat->Vh * prec_cbytes(p));
  *norm_off = data;
  return data + (p == P16 ? round256((size_t)n * lat->Vh * sizeof(float)) : 0);
}

DevBatch view(Lattice *lat, int p, int n, char *base, size_t norm_off) {
  DevBatch b;
  b.lat = lat;
  b.prec = p;
  b.n = n;
  b.data = base;
  b.norm = p == P16 ? (float *)(base + norm_off) : nullptr;
  return b;
}
}  // namespace

// 0: done; -2: the batches need more than the scratch array holds, or the
// links in the requested precision have not been built yet (nothing ran).
static int hop_noalloc_impl(QOP_D3_FermionLinksAsqtad *fla, int prec, int n, QOP_D3_ColorVector *out[],
                            QOP_D3_ColorVector *in[], int eo_out) {
  Links &L = fla->L;
  Lattice *lat = L.lat;
  QOP_REQUIRE(prec == 64 || prec == 32 || prec == 16, "hop_noalloc: precision must be 64, 32 or 16");
  QOP_REQUIRE(n >= 1 && n <= 64, "hop_noalloc: 1..64 vectors");
  QOP_REQUIRE(eo_out == QOP_EVEN || eo_out == QOP_ODD, "hop_noalloc: output subset must be EVEN or ODD");
  const int p = prec == 64 ? P64 : prec == 32 ? P32 : P16;
  size_t noff = 0;
  const size_t one = batch_bytes(lat, p, n, &noff);
  if (2 * one > kNoallocScratchBytes) {
    fprintf(stderr,
            "QOP-MI355X: QOPX_D3_asqtad_hop_noalloc: %d vectors of %ld sites in precision %d need %zu "
            "bytes of scratch, above its %zu\n",
            n, (long)lat->Vh, prec, 2 * one, kNoallocScratchBytes);
    return -2;
  }
  LinkSet *ls = p == P64 ? &L.l64 : p == P32 ? &L.l32 : &L.l16;
  if (!ls->data) {
    fprintf(stderr,
            "QOP-MI355X: QOPX_D3_asqtad_hop_noalloc: links in precision %d are not built yet (one "
            "QOPX_D3_asqtad_hop call in that precision builds them)\n",
            prec);
    return -2;
  }
  char *base = (char *)scratch_ptr();
  DevBatch src = view(lat, p, n, base, noff);
  DevBatch dst = view(lat, p, n, base + one, noff);
  const int src_par = 1 - eo_out;
  hipStream_t st = stream();
  int pin[1] = {P64};
  for (int b = 0; b < n; b++) {  // FP64 -> stored precision (FP16: per-site norm)
    LincombArgs a;
    a.Vh = (int)lat->Vh;
    a.out = src.slot(b);
    a.v[0] = in[b]->v.half(src_par);
    a.cr[0] = 1.0;
    launch_lincomb(a, p, pin, 1, false, st);
  }
  HopShape hs;  // the tuned shape when one exists (tuning allocates), else a plain one
  if (!cached_hop_shape(L, p, n, RED_NONE, &hs)) hs = HopShape{1, 256, 1, 0, 0, 0, 0};
  HopArgs a;
  a.g = lat->g;
  a.parity = eo_out;
  a.links = ls->parity_data(lat, eo_out);
  a.lscale = ls->parity_scale(lat, eo_out);
  a.src = src.slot(0);
  a.dst = dst.slot(0);
  a.sstride = a.dstride = 3 * lat->Vh;
  a.nstride = lat->Vh;
  a.n = n;
  hs.apply(a);
  launch_hop_args(a, p, ls->nl, hs.nb, hs.block, RED_NONE, false, st);
  int pd[1] = {p};
  for (int b = 0; b < n; b++) {  // stored precision -> FP64
    LincombArgs c;
    c.Vh = (int)lat->Vh;
    c.out = out[b]->v.half(eo_out);
    c.v[0] = dst.slot(b);
    c.cr[0] = 1.0;
    launch_lincomb(c, P64, pd, 1, false, st);
  }
  QOP_CHECK_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace qop

using namespace qop;

extern "C" {

// QOPX_D3_asqtad_hop without device allocation: see the file comment.
// Returns 0, or -2 when the request does not fit the scratch array or the
// links in precision `prec` have not been built (nothing is computed then).
int QOPX_D3_asqtad_hop_noalloc(QOP_D3_FermionLinksAsqtad *fla, int prec, int n, QOP_D3_ColorVector *out[],
                               QOP_D3_ColorVector *in[], QOP_evenodd_t eo_out) {
  int rc = -2;
  if (!guard_oom(__func__, nullptr, [&] { rc = hop_noalloc_impl(fla, prec, n, out, in, eo_out); })) return -2;
  return rc;
}

}  // extern "C"
