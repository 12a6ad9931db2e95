"""GPU parity of the Dslash and the link preparation against the oracle,
through the C ABI (QOP_D3_/QOP_F3_ entry points).

Tolerances (relative L2 over the written subset):
  FP64 Dslash <= 1e-13, FP32 Dslash <= 1e-6 (vs the FP64 oracle on FP32-rounded
  inputs), links/fattening FP64 <= 1e-13.
"""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC, from_raw

pytestmark = pytest.mark.gpu
EVEN, ODD, EVENODD = 0, 1, 2


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def setup(oracle):
    L = [4, 4, 6, 8]
    U = q.gen_gauge(L, 1001, 0.3)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    s = q.Session(L)
    fla = s.links_from_raw(fat, lng)
    h = oracle.links(L, fat, lng)
    yield dict(L=L, U=U, fat=fat, lng=lng, s=s, fla=fla, h=h)
    s.lib.QOP_D3_asqtad_destroy_L(fla)
    oracle.free_links(h)
    s.close()


@pytest.mark.parametrize("eo_out,eo_in", [(ODD, EVEN), (EVEN, ODD), (EVENODD, EVENODD),
                                          (EVEN, EVEN), (ODD, ODD), (EVEN, EVENODD),
                                          (ODD, EVENODD), (EVENODD, EVEN), (EVENODD, ODD)])
def test_dslash_fp64_all_subsets(setup, oracle, eo_out, eo_in):
    s, L = setup["s"], setup["L"]
    V = s.V
    src = q.gen_source(L, 2001, EVENODD)
    init = q.gen_source(L, 2002, EVENODD)  # sites outside eo_out must be untouched
    vin, vout = s.vec(src), s.vec(init)
    info = q.QOPInfo()
    s.lib.QOP_D3_asqtad_dslash(info, setup["fla"], 0.13, vout, vin, eo_out, eo_in)
    got = s.get(vout)
    ref = oracle.dslash(setup["h"], 0.13, src, eo_out, eo_in, out=init.copy())
    assert rel(got, ref) <= 1e-13
    sl = {EVEN: slice(V // 2, V), ODD: slice(0, V // 2), EVENODD: slice(0, 0)}[eo_out]
    assert np.array_equal(got[sl], init[sl])
    s.free(vin)
    s.free(vout)


def test_dslash_fp32(setup, oracle):
    s, L = setup["s"], setup["L"]
    V = s.V
    fat32 = [f.astype(np.float32) for f in setup["fat"]]
    lng32 = [x.astype(np.float32) for x in setup["lng"]]
    flf = s.links_from_raw(fat32, lng32, prec="F")
    src = q.gen_source(L, 2003, EVEN).astype(np.float32)
    vin, vout = s.vec(src, prec="F"), s.zeros(prec="F")
    s.lib.QOP_F3_asqtad_dslash(q.QOPInfo(), flf, 0.0, vout, vin, ODD, EVEN)
    got = s.get(vout, prec="F").astype(np.float64)
    h32 = oracle.links(L, [f.astype(np.float64) for f in fat32], [x.astype(np.float64) for x in lng32])
    ref = oracle.dslash(h32, 0.0, src.astype(np.float64), ODD, EVEN)
    assert rel(got[V // 2:], ref[V // 2:]) <= 1e-6
    # FD conversion of the double links gives the same FP32 operator
    flf2 = s.lib.QOP_FD3_asqtad_create_L_from_L(setup["fla"])
    vout2 = s.zeros(prec="F")
    s.lib.QOP_F3_asqtad_dslash(q.QOPInfo(), flf2, 0.0, vout2, vin, ODD, EVEN)
    got2 = s.get(vout2, prec="F")
    assert rel(got2[V // 2:].astype(np.float64), ref[V // 2:]) <= 1e-6
    oracle.free_links(h32)
    for v in (vin, vout, vout2):
        s.free(v, prec="F")
    s.lib.QOP_F3_asqtad_destroy_L(flf)
    s.lib.QOP_F3_asqtad_destroy_L(flf2)


def test_dslash_without_long_links(setup, oracle):
    s, L = setup["s"], setup["L"]
    fla8 = s.links_from_raw(setup["fat"], None)
    h8 = oracle.links(L, setup["fat"], None)
    src = q.gen_source(L, 2004, EVENODD)
    vin, vout = s.vec(src), s.zeros()
    s.lib.QOP_D3_asqtad_dslash(q.QOPInfo(), fla8, 0.2, vout, vin, EVENODD, EVENODD)
    ref = oracle.dslash(h8, 0.2, src, EVENODD, EVENODD)
    assert rel(s.get(vout), ref) <= 1e-13
    s.free(vin)
    s.free(vout)
    s.lib.QOP_D3_asqtad_destroy_L(fla8)
    oracle.free_links(h8)


def test_links_roundtrip_and_diaginv(setup):
    s = setup["s"]
    fat, lng = s.extract_links(setup["fla"])
    for mu in range(4):
        assert np.array_equal(fat[mu], setup["fat"][mu])
        assert np.array_equal(lng[mu], setup["lng"][mu])
    src = q.gen_source(setup["L"], 2005, EVENODD)
    vin, vout = s.vec(src), s.zeros()
    s.lib.QOP_D3_asqtad_diaginv(q.QOPInfo(), setup["fla"], 0.25, vout, vin, ODD)
    got = s.get(vout)
    V = s.V
    np.testing.assert_allclose(got[V // 2:], src[V // 2:] / 0.25, rtol=1e-15)
    assert np.all(got[: V // 2] == 0)
    s.free(vin)
    s.free(vout)


@pytest.mark.parametrize("with_long", [True, False])
def test_single_stored_style_matches_double_stored(setup, oracle, with_long):
    """Style 0 (single-stored links, backward links formed as -U(x-d)^dagger
    in the hop; lib/asqtad_dslash_p.c:1195-1244) gives the same operator bit
    for bit as style 1 (double-stored), in FP64 and FP32, and the same solve."""
    s, L = setup["s"], setup["L"]
    V = s.V
    lng = setup["lng"] if with_long else None
    fla = s.links_from_raw(setup["fat"], lng)
    h = oracle.links(L, setup["fat"], lng)
    src = q.gen_source(L, 2006, EVENODD)
    res = {}
    try:
        for st in (1, 0):
            assert s.set_opts(st=st) == 0
            vin, vout = s.vec(src), s.zeros()
            s.lib.QOP_D3_asqtad_dslash(q.QOPInfo(), fla, 0.3, vout, vin, EVENODD, EVENODD)
            res[st] = s.get(vout)
            s.free(vin)
            s.free(vout)
            flf = s.lib.QOP_FD3_asqtad_create_L_from_L(fla)
            vin, vout = s.vec(src.astype(np.float32), prec="F"), s.zeros(prec="F")
            s.lib.QOP_F3_asqtad_dslash(q.QOPInfo(), flf, 0.0, vout, vin, ODD, EVEN)
            res[(st, "F")] = s.get(vout, prec="F")
            s.free(vin, prec="F")
            s.free(vout, prec="F")
            s.lib.QOP_F3_asqtad_destroy_L(flf)
            b = src.copy()
            b[V // 2:] = 0
            ia = q.invert_arg_default()
            ia.evenodd = EVEN
            ra = q.resid_arg_default()
            ra.rsqmin = 1e-20
            vin, vout = s.vec(b), s.zeros()
            s.lib.QOP_D3_asqtad_invert(C.byref(q.QOPInfo()), fla, C.byref(ia), C.byref(ra), 0.1, vout, vin)
            res[(st, "x")] = (s.get(vout), ra.final_iter)
            s.free(vin)
            s.free(vout)
        assert rel(res[0], oracle.dslash(h, 0.3, src, EVENODD, EVENODD)) <= 1e-13
        assert np.array_equal(res[0], res[1])
        assert np.array_equal(res[(0, "F")], res[(1, "F")])
        assert np.array_equal(res[(0, "x")][0], res[(1, "x")][0])
        assert res[(0, "x")][1] == res[(1, "x")][1]
    finally:
        s.set_opts(st=1)
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        oracle.free_links(h)
