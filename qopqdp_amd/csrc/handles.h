// Definitions of the opaque QOP handle types of include/qop_mi355x.h.
#pragma once
#include <cstdio>
#include <stdexcept>

#include "../../include/qop_mi355x.h"
#include "common.h"

struct QOP_D3_ColorVector_struct {
  qop::DevVec v;
  double *raw = nullptr;  // set by convert_V_from_raw
};
struct QOP_F3_ColorVector_struct {
  qop::DevVec v;
  float *raw = nullptr;
};
// Gauge fields stay on the device in the raw layout (FP64), one array per
// direction.
// raw: the caller's arrays kept by convert_*_from_raw (written back and
// returned by convert_*_to_raw).
struct QOP_D3_GaugeField_struct {
  qop::Lattice *lat = nullptr;
  double *U[4] = {nullptr, nullptr, nullptr, nullptr};
  double **raw = nullptr;
};
struct QOP_F3_GaugeField_struct {
  qop::Lattice *lat = nullptr;
  double *U[4] = {nullptr, nullptr, nullptr, nullptr};
  float **raw = nullptr;
};
// Force fields: device FP64, raw layout, one array per direction.
struct QOP_D3_Force_struct {
  qop::Lattice *lat = nullptr;
  double *F[4] = {nullptr, nullptr, nullptr, nullptr};
  double **raw = nullptr;
};
struct QOP_F3_Force_struct {
  qop::Lattice *lat = nullptr;
  double *F[4] = {nullptr, nullptr, nullptr, nullptr};
  float **raw = nullptr;
};
// raw_fat / raw_lng: the caller's link arrays taken by convert_L_from_raw
// (they hold the links x 1/2 until convert_L_to_raw).
struct QOP_D3_FermionLinksAsqtad_struct {
  qop::Links L;
  double **raw_fat = nullptr, **raw_lng = nullptr;
};
struct QOP_F3_FermionLinksAsqtad_struct {
  qop::Links L;
  float **raw_fat = nullptr, **raw_lng = nullptr;
};

namespace qop {
// Link preparation shared by the QOP entry points and the request server
// (api.cpp).  links_from_raw64: FP64 raw fat/long links (host, already
// rephased and smeared, lng may be null).  links_from_gauge64: raw FP64 gauge
// field (host) -> device, rephase (r0/bc/sign as QOP_D3_rephase_G; null
// skips), asqtad smearing with coefficients c.
void links_from_raw64(Links &L, double *fat[4], double *lng[4]);
void links_from_gauge64(Links &L, double *const U_host[4], const QOP_asqtad_coeffs_t *c, int *r0,
                        QOP_bc_t *bc, QOP_staggered_sign_t *sign);

// Runs body(); a device allocation failure inside it is reported on stderr
// and, when info is given, as info->status = QOP_MEM_ERROR (the QOP status
// for it, include/qop_int.h:19) instead of ending the process; an argument
// the device code cannot take (std::runtime_error, e.g. a lattice beyond the
// kernels' 32-bit site offsets) as QOP_FAIL.  Objects built inside release
// their memory while the failure unwinds.  Returns false on such a failure.
template <class F>
bool guard_oom(const char *where, QOP_info_t *info, F &&body) {
  try {
    body();
    return true;
  } catch (const DeviceOutOfMemory &e) {
    fprintf(stderr, "QOP-MI355X: %s: device out of memory (an allocation of %zu bytes failed)\n",
            where, e.bytes);
    fflush(stderr);
    if (info) info->status = QOP_MEM_ERROR;
    return false;
  } catch (const std::runtime_error &e) {
    fprintf(stderr, "QOP-MI355X: %s: %s\n", where, e.what());
    fflush(stderr);
    if (info) info->status = QOP_FAIL;
    return false;
  }
}
}  // namespace qop
