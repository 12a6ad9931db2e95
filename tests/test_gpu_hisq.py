"""HISQ link preparation on the GPU (QOP_D3_hisq_create_L_from_G) against the
oracle chain fat7 -> U(3) projection -> asqtad level 2 (oracle/hisq_oracle.c,
lib/hisq_links_p.c:62-256), the forced-fallback projection, the epsilon-Naik
link sets, the free field, and mixed-precision solves on the HISQ links.

Tolerances: links relative L2 <= 1e-12 (FP64, different summation order);
fallback projection (Newton polar iteration against the oracle's analytic
formula) <= 1e-11; solves: FP64 true residual <= tol."""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from test_oracle_hisq import HISQ_FAT7, HISQ_L2

pytestmark = pytest.mark.gpu
EVEN, ODD, EVENODD = 0, 1, 2


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_hisq(s, U, coeffs, opts=None):
    g = s.gauge(U)
    r0, bc, sign, keep = q.milc_rephase_args(s.L)
    s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
    if opts:
        arr = (q.QOPOpt * len(opts))()
        names = [k.encode() for k in opts]
        for i, (k, v) in enumerate(opts.items()):
            arr[i] = q.QOPOpt(names[i], float(v), None)
        assert s.lib.QOP_hisq_links_set_opts(arr, len(opts)) == 0
    info = q.QOPInfo()
    h = s.lib.QOP_D3_hisq_create_L_from_G(C.byref(info), C.byref(coeffs), g)
    s.lib.QOP_D3_destroy_G(g)
    return h, info


def reset_opts(s):
    names = [b"reunit_svd_only", b"want_deps"]
    arr = (q.QOPOpt * 2)(q.QOPOpt(names[0], 0.0, None), q.QOPOpt(names[1], 0.0, None))
    s.lib.QOP_hisq_links_set_opts(arr, 2)


def oracle_hisq(oracle, L, U, l2=HISQ_L2):
    Ur = [u.copy() for u in U]
    oracle.rephase_milc(L, Ur)
    W, nsvd, _ = oracle.hisq_W(L, Ur, HISQ_FAT7)
    assert nsvd == 0
    return W, oracle.make_imp_links(L, W, l2)


def test_hisq_links_match_oracle(oracle):
    L = [4, 4, 4, 8]
    U = q.gen_gauge(L, 4001, 0.3)
    s = q.Session(L)
    try:
        h, info = gpu_hisq(s, U, q.hisq_coeffs())
        assert info.count1 == 0  # no fallback projection
        fla = s.lib.QOP_D3_get_asqtad_links_from_hisq(h)[0]
        fat, lng = s.extract_links(fla)
        _, (ofat, olng) = oracle_hisq(oracle, L, U)
        for mu in range(4):
            assert rel(fat[mu], ofat[mu]) <= 1e-12
            assert rel(lng[mu], olng[mu]) <= 1e-12
        assert s.lib.QOP_D3_get_asqtad_deps_links_from_hisq(h) is None
        s.lib.QOP_D3_hisq_destroy_L(h)
    finally:
        s.close()


def test_hisq_forced_fallback_projection(oracle):
    L = [4, 4, 4, 8]
    V = int(np.prod(L))
    U = q.gen_gauge(L, 4002, 0.3)
    s = q.Session(L)
    try:
        h, info = gpu_hisq(s, U, q.hisq_coeffs(), {"reunit_svd_only": 1})
        assert info.count1 == 4 * V
        fat, lng = s.extract_links(s.lib.QOP_D3_get_asqtad_links_from_hisq(h)[0])
        _, (ofat, olng) = oracle_hisq(oracle, L, U)
        for mu in range(4):
            assert rel(fat[mu], ofat[mu]) <= 1e-11
            assert rel(lng[mu], olng[mu]) <= 1e-11
        s.lib.QOP_D3_hisq_destroy_L(h)
    finally:
        reset_opts(s)
        s.close()


def test_hisq_eps_naik_sets(oracle):
    """n_naiks = 2: fn[0] = level-2 links, fn[1] = level-2 + eps x difference
    links; want_deps keeps the difference links."""
    L = [4, 4, 4, 8]
    U = q.gen_gauge(L, 4003, 0.3)
    diff = dict(one_link=0.05, three_staple=0, five_staple=0, seven_staple=0, lepage=0,
                naik=-1 / 24)
    s = q.Session(L)
    try:
        c = q.hisq_coeffs(eps_naik=(0.0, 0.3), difference_one_link=0.05, difference_naik=-1 / 24)
        h, info = gpu_hisq(s, U, c, {"want_deps": 1})
        fn = s.lib.QOP_D3_get_asqtad_links_from_hisq(h)
        W, (f2, l2) = oracle_hisq(oracle, L, U)
        f3, l3 = oracle.make_imp_links(L, W, diff)
        fat0, lng0 = s.extract_links(fn[0])
        fat1, lng1 = s.extract_links(fn[1])
        fatd, lngd = s.extract_links(s.lib.QOP_D3_get_asqtad_deps_links_from_hisq(h))
        for mu in range(4):
            assert rel(fat0[mu], f2[mu]) <= 1e-12 and rel(lng0[mu], l2[mu]) <= 1e-12
            assert rel(fat1[mu], f2[mu] + 0.3 * f3[mu]) <= 1e-12
            assert rel(lng1[mu], l2[mu] + 0.3 * l3[mu]) <= 1e-12
            assert rel(fatd[mu], f3[mu]) <= 1e-12 and rel(lngd[mu], l3[mu]) <= 1e-12
        s.lib.QOP_D3_hisq_destroy_L(h)
    finally:
        reset_opts(s)
        s.close()


def test_hisq_free_field_spectrum():
    """U = 1: fat7 weight 1, W = U, level-2 weights 9/8 and -1/24: the even
    operator has the asqtad free spectrum (GPU links, GPU Dslash)."""
    L = [4, 4, 4, 8]
    V = int(np.prod(L))
    eye = np.zeros((V, 3, 3, 2))
    eye[:, [0, 1, 2], [0, 1, 2], 0] = 1
    s = q.Session(L)
    try:
        h, _ = gpu_hisq(s, [eye] * 4, q.hisq_coeffs())
        fla = s.lib.QOP_D3_get_asqtad_links_from_hisq(h)[0]
        x = q.raw_coords_all(L)
        m = 0.05
        for n in ([0, 0, 0, 0], [1, 2, 0, 3], [3, 3, 1, 6]):
            p = np.array([2 * np.pi * n[i] / L[i] for i in range(3)] + [(2 * n[3] + 1) * np.pi / L[3]])
            v = np.exp(1j * (x @ p))[:, None] * np.array([1, 0, 0])[None, :]
            v[V // 2:] = 0
            raw = np.ascontiguousarray(np.stack([v.real, v.imag], -1))
            vin, t, out = s.vec(raw), s.zeros(), s.zeros()
            info = q.QOPInfo()
            s.lib.QOP_D3_asqtad_dslash(info, fla, 0.0, t, vin, ODD, EVEN)
            s.lib.QOP_D3_asqtad_dslash(info, fla, 0.0, out, t, EVEN, ODD)
            Av = m * m * raw - s.get(out)
            sv = 9 / 8 * np.sin(p) - 1 / 24 * np.sin(3 * p)
            lam = m * m + np.sum(sv * sv)
            np.testing.assert_allclose(Av[: V // 2], lam * raw[: V // 2], atol=1e-12)
            for vv in (vin, t, out):
                s.free(vv)
        s.lib.QOP_D3_hisq_destroy_L(h)
    finally:
        s.close()


@pytest.mark.parametrize("prec", [16, 32])
def test_mixed_solves_on_hisq_links(oracle, prec):
    """Batched multi-mass solves on HISQ links (FP16 links with per-link
    scales for prec 16): every mass to FP64 true residual 1e-10."""
    from test_gpu_invert import batch_api
    L = [4, 4, 4, 8]
    U = q.gen_gauge(L, 4004, 0.3)
    s = q.Session(L)
    try:
        h, _ = gpu_hisq(s, U, q.hisq_coeffs())
        fla = s.lib.QOP_D3_get_asqtad_links_from_hisq(h)[0]
        _, (ofat, olng) = oracle_hisq(oracle, L, U)
        oh = oracle.links(L, ofat, olng)
        srcs = [q.gen_source(L, 4100 + i, EVEN) for i in range(2)]
        masses = [0.002, 0.0246038, 0.5]
        ok, outs, ras = batch_api(s, fla, srcs, masses, 1e-10, prec, multi=True, capacity=6)
        assert ok == len(outs)
        for i, b in enumerate(srcs):
            for j, m in enumerate(masses):
                assert oracle.true_resid_even(oh, m, outs[i * len(masses) + j], b) <= 1e-10
        oracle.free_links(oh)
        s.lib.QOP_D3_hisq_destroy_L(h)
    finally:
        s.close()
