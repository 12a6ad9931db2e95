"""GPU eigCG (option cg = 1: the inner solver of the EVENODD invert,
lib/asqtad_invert_p.c:683-687, 767-769; lib/invert_eigcg_p.c) through the C
ABI, against the oracle's eigCG (tests/test_oracle_eigcg.py pins it against
dense eigenvalues) on the same sequence of sources: every solution equals
the oracle's to 1e-8 relative and reaches the residual; the iteration counts
agree within 2 and fall as the deflation space grows.  Option checks: sizes
the restart cannot use are refused.  The deflation space lives in the links
object and is freed with it."""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC, gauge_numpy

pytestmark = pytest.mark.gpu
EVEN, ODD, EVENODD = 0, 1, 2
L = [4, 4, 4, 8]
V = 512
MASS = 0.003


def set_opts(s, **kw):
    return s.set_opts(**kw)


def sources(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        b = rng.standard_normal((V, 3, 2))
        b[V // 2:] *= 0.5  # the even part dominates: the CG runs on even sites
        out.append(b)
    return out


def solve(s, fla, b, rsqmin):
    ia = q.invert_arg_default()
    ia.evenodd = EVENODD
    ra = q.resid_arg_default()
    ra.rsqmin = rsqmin
    info = q.QOPInfo()
    vin = s.vec(b)
    vout = s.zeros()
    s.lib.QOP_D3_asqtad_invert(C.byref(info), fla, C.byref(ia), C.byref(ra), MASS, vout, vin)
    x = s.get(vout)
    s.free(vin)
    s.free(vout)
    assert info.status == 0
    return x, ra


def test_eigcg_matches_oracle(oracle):
    U = gauge_numpy(L, 11, 0.3)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    h = oracle.links(L, fat, lng)
    eg = oracle.eigcg_state()
    s = q.Session(L)
    try:
        fla = s.links_from_raw(fat, lng)
        assert set_opts(s, cg=1, eigcg_nev=6, eigcg_m=24, eigcg_l=100) == 0
        iters = []
        for b in sources(1, 8):
            x, ra = solve(s, fla, b, 1e-20)
            xo, rao = oracle.invert_eigcg(h, eg, MASS, b, 1e-20, numax=100, m=24, nev=6)
            assert ra.final_rsq <= 1e-20
            assert np.abs(x - xo).max() <= 1e-8 * np.abs(xo).max()
            assert abs(ra.final_iter - rao.final_iter) <= 2, (ra.final_iter, rao.final_iter)
            iters.append(ra.final_iter)
        assert iters[-1] <= 0.9 * iters[0], iters
        s.lib.QOP_D3_asqtad_destroy_L(fla)  # frees the deflation space
    finally:
        set_opts(s, cg=0)
        s.close()
        oracle.eigcg_free(eg)
        oracle.free_links(h)


def test_eigcg_option_checks():
    s = q.Session(L)
    try:
        assert set_opts(s, cg=1, eigcg_nev=6, eigcg_m=12, eigcg_l=100) != 0  # m < 2 nev + 2
        assert set_opts(s, cg=1, eigcg_nev=4, eigcg_m=24, eigcg_l=20) != 0   # l < m
        assert set_opts(s, cg=1, eigcg_nev=0, eigcg_m=24, eigcg_l=100) != 0
        assert set_opts(s, cg=1, eigcg_nev=4, eigcg_m=24, eigcg_l=100) == 0
    finally:
        set_opts(s, cg=0, eigcg_nev=10, eigcg_m=100, eigcg_l=200)
        s.close()


def test_eigcg_larger_lattice_deflates():
    """16^3x32 at m = 0.01: every solve to 1e-16 (true full-D residual), the
    later solves take fewer iterations than the first."""
    L2 = [16, 16, 16, 32]
    s = q.Session(L2)
    try:
        g = s.gauge(q.gen_gauge(L2, 12, 0.3))
        r0, bc, sign, keep = q.milc_rephase_args(L2)
        s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
        fla, info = s.links_from_gauge(g)
        assert info.status == 0
        assert set_opts(s, cg=1, eigcg_nev=8, eigcg_m=40, eigcg_l=120) == 0
        iters = []
        for k in range(5):
            b = q.gen_source(L2, 13 + k, EVEN)
            ia = q.invert_arg_default()
            ia.evenodd = EVENODD
            ra = q.resid_arg_default()
            ra.rsqmin = 1e-16
            inf = q.QOPInfo()
            vin = s.vec(b)
            vout = s.zeros()
            s.lib.QOP_D3_asqtad_invert(C.byref(inf), fla, C.byref(ia), C.byref(ra), 0.01, vout, vin)
            assert inf.status == 0 and ra.final_rsq <= 1e-16
            x = s.get(vout)
            assert np.isfinite(x).all()
            s.free(vin)
            s.free(vout)
            iters.append(ra.final_iter)
        assert iters[-1] < iters[0], iters
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        s.lib.QOP_D3_destroy_G(g)
    finally:
        set_opts(s, cg=0, eigcg_nev=10, eigcg_m=100, eigcg_l=200)
        s.close()
