"""QOPX_D3_asqtad_hop_noalloc (staging batches in the library's fixed device
array instead of allocated memory) gives the same bits as QOPX_D3_asqtad_hop
in every stored precision, on two calls in a row whose second has fewer
vectors; it refuses requests that do not fit the array and precisions whose
links are not built, leaving the outputs untouched."""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC

pytestmark = pytest.mark.gpu
EVEN, ODD = 0, 1
P = C.c_void_p


def _links(s, oracle, L, seed):
    U = q.gen_gauge(L, seed, 0.3)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    return s.links_from_raw(fat, lng)


def _call(fn, fla, prec, srcs, eo_out, V, fill=0.0):
    n = len(srcs)
    vin = [q_s.vec(x) for x in srcs]
    vout = [q_s.vec(np.full((V, 3, 2), fill)) for _ in range(n)]
    rc = fn(fla, prec, n, (P * n)(*vout), (P * n)(*vin), eo_out)
    out = [q_s.get(v) for v in vout]
    for v in vin + vout:
        q_s.free(v)
    return rc, out


q_s = None


def test_noalloc_hop_matches_hop(oracle):
    global q_s
    L = [8, 8, 8, 8]
    q_s = s = q.Session(L)
    try:
        fla = _links(s, oracle, L, 1300)
        V = s.V
        srcs = [q.gen_source(L, 1400 + i, EVEN) for i in range(6)]
        # FP16 links are built by the first FP16 hop: until then, refused
        rc, out = _call(s.lib.QOPX_D3_asqtad_hop_noalloc, fla, 16, srcs[:2], ODD, V, fill=7.0)
        assert rc == -2
        assert all(np.all(o == 7.0) for o in out)
        for prec in (64, 32, 16):
            for eo_out, batch in ((ODD, srcs), (ODD, srcs[3:5]), (EVEN, [q.gen_source(L, 1500, ODD)])):
                ref_rc, ref = _call(s.lib.QOPX_D3_asqtad_hop, fla, prec, batch, eo_out, V)
                assert ref_rc == 0
                rc, got = _call(s.lib.QOPX_D3_asqtad_hop_noalloc, fla, prec, batch, eo_out, V)
                assert rc == 0, (prec, len(batch))
                sl = slice(V // 2, V) if eo_out == ODD else slice(0, V // 2)
                for g, r in zip(got, ref):
                    assert np.array_equal(g[sl], r[sl]), (prec, len(batch), eo_out)
                    assert np.any(g[sl] != 0)
        s.lib.QOP_D3_asqtad_destroy_L(fla)
    finally:
        s.close()
        q_s = None


def test_noalloc_hop_refuses_what_does_not_fit(oracle):
    """16^4: 32 FP64 vectors in and out need 2 x 32 x 48 B x 32768 sites =
    96 MiB of staging, above the 64 MiB array; 16 need 48 MiB and run."""
    global q_s
    L = [16, 16, 16, 16]
    q_s = s = q.Session(L)
    try:
        fla = _links(s, oracle, L, 1310)
        V = s.V
        srcs = [q.gen_source(L, 1600 + i, EVEN) for i in range(32)]
        rc, out = _call(s.lib.QOPX_D3_asqtad_hop_noalloc, fla, 64, srcs, ODD, V, fill=-3.0)
        assert rc == -2
        assert all(np.all(o == -3.0) for o in out[:4])
        ref_rc, ref = _call(s.lib.QOPX_D3_asqtad_hop, fla, 64, srcs[:16], ODD, V)
        assert ref_rc == 0
        rc, got = _call(s.lib.QOPX_D3_asqtad_hop_noalloc, fla, 64, srcs[:16], ODD, V)
        assert rc == 0
        for g, r in zip(got, ref):
            assert np.array_equal(g[V // 2:], r[V // 2:])
        s.lib.QOP_D3_asqtad_destroy_L(fla)
    finally:
        s.close()
        q_s = None
