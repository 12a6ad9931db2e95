// Internal declarations shared by the MI355X asqtad library.
//
// Device data layout (one lattice, volume V, Vh = V/2):
//   site order   : QOP raw order (lib/layout_user.c:237-258): lexicographic
//                  l = ((x0*L1+x1)*L2+x2)*L3+x3, even sites [0,Vh), odd
//                  sites [Vh,V); within a parity, h = l/2.
//   colour vector: [parity][colour 3][h] complex  (structure of arrays,
//                  coalesced over h).  FP16 vectors add a float norm[h] per
//                  parity: stored value = v / norm, norm = max |component|.
//   links        : double stored per output parity p:
//                  [p][k 0..nl-1][element 0..8][h] complex, k = 4*mu + j with
//                  j = {0: F(x), 1: L(x), 2: -F(x-mu)^+, 3: -L(x-3mu)^+}, all
//                  already multiplied by 1/2 (lib/asqtad_dslash_p.c:98-104,
//                  123-127, 201-206).  Without long links nl = 8 and
//                  k = 2*mu + {0: F(x), 1: -F(x-mu)^+}.  FP16 links add a
//                  float scale per link [p][k][h] and store each element
//                  as the half pair (re, -im) / scale.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace qop {

// P32H is a hop precision only: FP32 vectors with the FP16 links (+ per-link
// scale) of the P16 link set.
enum Prec { P64 = 0, P32 = 1, P16 = 2, P32H = 3 };
inline size_t prec_cbytes(int p) { return p == P64 ? 16 : p == P16 ? 4 : 8; }

[[noreturn]] void fatal(const char *file, int line, const char *msg);
#define QOP_CHECK_HIP(x)                                                   \
  do {                                                                     \
    hipError_t _e = (x);                                                   \
    if (_e != hipSuccess) ::qop::fatal(__FILE__, __LINE__, hipGetErrorString(_e)); \
  } while (0)
#define QOP_REQUIRE(c, msg)                                                \
  do {                                                                     \
    if (!(c)) ::qop::fatal(__FILE__, __LINE__, msg);                       \
  } while (0)

// Device memory exhausted (hipErrorOutOfMemory).  Thrown by dev_alloc; the
// API entry points catch it, release what they allocated and report
// QOP_MEM_ERROR (or return null) instead of exiting.
struct DeviceOutOfMemory {
  size_t bytes;
};
void dev_alloc_bytes(void **p, size_t bytes);
template <class T>
inline void dev_alloc(T **p, size_t bytes) {
  dev_alloc_bytes((void **)p, bytes);
}
// Owning device buffer for temporaries: freed when it goes out of scope,
// also while a DeviceOutOfMemory unwinds.
template <class T>
struct DevBuf {
  T *p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { dev_alloc(&p, bytes); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  T *release() {
    T *q = p;
    p = nullptr;
    return q;
  }
};

// A stream-captured hipGraph and its executable instance, owned together:
// the captured graph lives as long as the instance made from it and both are
// destroyed in one place (instance first), also while an exception unwinds.
struct CapturedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  CapturedGraph() = default;
  CapturedGraph(const CapturedGraph &) = delete;
  CapturedGraph &operator=(const CapturedGraph &) = delete;
  ~CapturedGraph() { reset(); }
  explicit operator bool() const { return exec != nullptr; }
  void begin(hipStream_t st) {
    reset();
    QOP_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  }
  void end(hipStream_t st) {
    QOP_CHECK_HIP(hipStreamEndCapture(st, &graph));
    QOP_CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  void launch(hipStream_t st) const { QOP_CHECK_HIP(hipGraphLaunch(exec, st)); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

// Lattice geometry passed by value to kernels.
struct Geom {
  int L[4];
  int Lh3;      // L3/2
  int Vh;       // sites per parity
  int stride[4];  // lexicographic strides {L1L2L3, L2L3, L3, 1}
};

struct Lattice {
  int L[4];
  long V, Vh;
  Geom g;
};
Lattice *lattice_get(const int L[4]);  // cached, never freed

// A parity-half (or batch of halves) of colour vectors in device memory.
struct VecView {
  void *data;     // [3][Vh] complex of the precision
  float *norm;    // FP16 only: [Vh]
};

// Full-lattice device colour vector.
struct DevVec {
  Lattice *lat = nullptr;
  int prec = P64;
  void *data = nullptr;   // [2][3][Vh] complex
  float *norm = nullptr;  // FP16: [2][Vh]
  VecView half(int parity) const {
    VecView v;
    v.data = (char *)data + (size_t)parity * 3 * lat->Vh * prec_cbytes(prec);
    v.norm = norm ? norm + (size_t)parity * lat->Vh : nullptr;
    return v;
  }
};
void devvec_alloc(DevVec &v, Lattice *lat, int prec);
void devvec_free(DevVec &v);

// Batch of parity halves: slot b at data + b*3*Vh complex; norms at norm + b*Vh.
struct DevBatch {
  Lattice *lat = nullptr;
  int prec = P32;
  int n = 0;
  void *data = nullptr;
  float *norm = nullptr;
  VecView slot(int b) const {
    VecView v;
    v.data = (char *)data + (size_t)b * 3 * lat->Vh * prec_cbytes(prec);
    v.norm = norm ? norm + (size_t)b * lat->Vh : nullptr;
    return v;
  }
};
void devbatch_alloc(DevBatch &v, Lattice *lat, int prec, int n);
void devbatch_free(DevBatch &v);

// Double-stored links of one precision.
struct LinkSet {
  int prec = -1;
  int nl = 0;            // 16 with Naik, 8 without
  void *data = nullptr;  // [2][nl][9][Vh] complex
  float *scale = nullptr;  // FP16: [2][nl][Vh]
  const void *parity_data(const Lattice *lat, int p) const {
    return (const char *)data + (size_t)p * nl * 9 * lat->Vh * prec_cbytes(prec);
  }
  const float *parity_scale(const Lattice *lat, int p) const {
    return scale ? scale + (size_t)p * nl * lat->Vh : nullptr;
  }
};
void linkset_free(LinkSet &l);

// Link object behind QOP_{D,F}3_FermionLinksAsqtad.
struct EigcgState;  // eigcg.h

struct Links {
  Lattice *lat = nullptr;
  EigcgState *eig = nullptr;  // eigCG deflation space (option cg = 1), owned
  int nl = 16;
  LinkSet l64, l32, l16;
  // single-stored forward links (styles 0 and 2: backward links are formed
  // in the hop as -U(x-d)^dagger): [p][nl/2][9][Vh], FP64 / FP32
  LinkSet s64, s32;
  LinkSet &get(int prec);         // builds lower precisions from l64 on demand
  LinkSet &get_single(int prec);  // builds the single-stored set on demand
  void drop_derived();            // frees every set derived from l64
  void free_all();
};

// ---------------------------------------------------------------------
// runtime
hipStream_t stream();  // library stream (default priority)
void sync();

// Reduction scratch: partial sums per 64-site chunk, results in doubles.
double *red_partial(size_t ndoubles);  // device, grows on demand
double *red_result(size_t ndoubles);   // device
double *red_host(size_t ndoubles);     // pinned host

// Link preparation (links.hip)
void build_links64(Links &L, const double *fat_dev[4], const double *lng_dev[4],
                   hipStream_t st);
void convert_links(const Lattice *lat, const LinkSet &src, LinkSet &dst,
                   int prec, hipStream_t st);
void extract_links64(const Links &L, double *fat_dev[4], double *lng_dev[4],
                     hipStream_t st);
void launch_rephase(const Lattice *lat, double *U_dev, int mu, const int r0[4],
                    double bcre, double bcim, int bcdir, int signbits,
                    int nslice, int sign_rel, hipStream_t st);
void launch_fatten(const Lattice *lat, const double *gf[4], const double *gl[4],
                   const double coef[6], double *fat[4], double *lng[4],
                   hipStream_t st);

// Asqtad fermion force (force.hip): F[mu] (raw FP64, device) += -2 TA(U_mu
// deriv_mu^+) for sum_t eps[t] x_t x_t^+ (full-lattice FP64 or FP32 vectors);
// coef = {one_link, three_staple, five_staple, seven_staple, lepage, naik};
// flops (optional) receives the reference's flop count.
void asqtad_force_dev(const Lattice *lat, double *const U_raw[4], double *F_raw[4],
                      const double coef[6], const double *eps, const DevVec *const *xs, int n,
                      double *flops, hipStream_t st);

// Asqtad fattening of raw FP64 device links (force.hip): fat = smeared gf,
// lng (optional) = naik gl gl gl; coef as for asqtad_force_dev.
void fatten_dev(const Lattice *lat, const double *const gf[4], const double *const gl[4],
                const double coef[6], double *fat[4], double *lng[4], hipStream_t st);

// HISQ fermion force (force.hip): F[mu] += TA(U_mu deriv_mu^+) with deriv
// through level 2 (W, lev2 or the difference set diff per epsilon-Naik group),
// the U(3) projection of V and the fat7 level on U (raw FP64 device links);
// the vectors are grouped by n_orders[0..n_naiks-1]; filter: force filter;
// nfilt receives the number of filtered links.
void hisq_force_dev(const Lattice *lat, double *const U_raw[4], double *const V_raw[4],
                    double *const W_raw[4], double *F_raw[4], const double fat7[6],
                    const double lev2[6], const double diff[6], int n_naiks, const double *eps_naik,
                    const double *eps, const DevVec *const *xs, const int *n_orders, double filter,
                    double *flops, int *nfilt, hipStream_t st);

// raw <-> SoA (kernels.hip); eo selects the sites converted.
void launch_raw_to_vec(const Lattice *lat, const double *raw_dev, DevVec &v,
                       int eo, hipStream_t st);
void launch_vec_to_raw(const Lattice *lat, const DevVec &v, double *raw_dev,
                       int eo, hipStream_t st);
void launch_rawf_to_vec(const Lattice *lat, const float *raw_dev, DevVec &v,
                        int eo, hipStream_t st);
void launch_vec_to_rawf(const Lattice *lat, const DevVec &v, float *raw_dev,
                        int eo, hipStream_t st);

}  // namespace qop
