// Definitions of the opaque QOP handle types of include/qop_mi355x.h.
#pragma once
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
struct QOP_D3_GaugeField_struct {
  qop::Lattice *lat = nullptr;
  double *U[4] = {nullptr, nullptr, nullptr, nullptr};
};
struct QOP_F3_GaugeField_struct {
  qop::Lattice *lat = nullptr;
  double *U[4] = {nullptr, nullptr, nullptr, nullptr};
};
struct QOP_D3_FermionLinksAsqtad_struct {
  qop::Links L;
};
struct QOP_F3_FermionLinksAsqtad_struct {
  qop::Links L;
};
