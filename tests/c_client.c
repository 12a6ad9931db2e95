/* A QOP client written against the reference API's generic names, built
 * against include/qop_mi355x.h: an EVEN-subset asqtad solve from raw fat and
 * long links (the call sequence of examples/asqtad_test.c:240-330).  Compiled
 * and linked (not run) by the CPU tests; it runs only with a device. */
#define QOP_PrecisionInt 2
#include "qop_mi355x.h"

#include <stdio.h>
#include <stdlib.h>

int solve(QOP_Real *src_raw, QOP_Real *dst_raw, QOP_Real *fat[4], QOP_Real *lng[4], double mass,
          double rsqmin) {
  QOP_info_t info = QOP_INFO_ZERO;
  QOP_invert_arg_t inv = QOP_INVERT_ARG_DEFAULT;
  QOP_resid_arg_t res = QOP_RESID_ARG_DEFAULT;
  inv.evenodd = QOP_EVEN;
  res.rsqmin = rsqmin;
  QOP_FermionLinksAsqtad *fla = QOP_asqtad_create_L_from_raw(fat, lng, QOP_EVENODD);
  QOP_ColorVector *in = QOP_create_V_from_raw(src_raw, QOP_EVEN);
  QOP_ColorVector *out = QOP_create_V_from_raw(dst_raw, QOP_EVEN);
  QOP_asqtad_invert(&info, fla, &inv, &res, mass, out, in);
  QOP_extract_V_to_raw(dst_raw, out, QOP_EVEN);
  QOP_destroy_V(in);
  QOP_destroy_V(out);
  QOP_asqtad_destroy_L(fla);
  printf("iters %d rsq %g status %d\n", res.final_iter, res.final_rsq, (int)info.status);
  return info.status == QOP_SUCCESS ? 0 : 1;
}

int main(void) {
  int L[4] = {4, 4, 4, 8};
  int V = L[0] * L[1] * L[2] * L[3];
  QOP_layout_t lay = QOP_LAYOUT_ZERO;
  lay.latdim = 4;
  lay.latsize = L;
  lay.machdim = -1;
  lay.sites_on_node = V;
  if (!QOPX_device_ok()) {
    printf("no device\n");
    return 0;
  }
  if (QOP_init(&lay) != QOP_SUCCESS) return 1;
  QOP_Real *U[4], *fat[4], *lng[4];
  for (int mu = 0; mu < 4; mu++) {
    U[mu] = calloc((size_t)V * 18, sizeof(QOP_Real));
    fat[mu] = U[mu];
    lng[mu] = U[mu];
  }
  QOPX_gen_gauge(L, 1, 0.3, U);
  QOP_Real *src = calloc((size_t)V * 6, sizeof(QOP_Real));
  QOP_Real *dst = calloc((size_t)V * 6, sizeof(QOP_Real));
  QOPX_gen_gaussian_V(L, 2, QOP_EVEN, src);
  int rc = solve(src, dst, fat, lng, 0.1, 1e-16);
  QOP_finalize();
  return rc;
}
