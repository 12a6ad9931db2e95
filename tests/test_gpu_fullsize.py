"""Parity at the workloads' own lattice sizes, with the oracle's FP64
operator as the residual checker:

  * 32^3x64: the single-RHS and 16-RHS hop in FP32 (<= 1e-6 relative L2)
    and FP16 (<= 1e-3, the bound set by FP16 vectors with per-site norms and
    FP16 links with per-link scales: unit roundoff 2^-11 ~ 4.9e-4 per
    stored value) against the oracle's FP64 hop; 16 Gaussian sources solved
    by the FP16 batched CG at m = 0.01 to FP64 true relative residual 1e-10;
  * 24^3x48: 16 sources x 8 masses (0.005 .. 0.2) by the batched multi-mass
    solver, every (source, mass) to 1e-10;
  * 48^3x96: HISQ links (fat7, U(3) projection, level 2), one source x 12
    masses (0.002 .. 0.5) to 1e-10, the GPU's links extracted and the
    residuals computed by the oracle's operator on them;
  * a serving mix over 16^4, 24^3x48 and 32^3x64: each request alone and in
    the concurrent mix gives the same bits and reaches its tolerance.

Inputs follow the workloads' recipe: seeded reunit(exp(0.3 i H)) links, MILC
phases and antiperiodic time, asqtad (or HISQ) smearing, unit-variance
Gaussian even-site sources."""
import ctypes as C
import gc

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC

pytestmark = pytest.mark.gpu
EVEN, ODD, EVENODD = 0, 1, 2
P = C.c_void_p
L2 = [32, 32, 32, 64]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="class")
def c2(oracle):
    """32^3x64 gauge field, the oracle's fattened links and link object, and
    the library's links built from the same raw fat/long links."""
    U = q.gen_gauge(L2, 1000, 0.3)
    oracle.rephase_milc(L2, U)
    fat, lng = oracle.make_imp_links(L2, U, MILC)
    del U
    h = oracle.links(L2, fat, lng)
    s = q.Session(L2)
    fla = s.links_from_raw(fat, lng)
    del fat, lng
    gc.collect()
    yield dict(s=s, fla=fla, h=h)
    s.lib.QOP_D3_asqtad_destroy_L(fla)
    oracle.free_links(h)
    s.close()


def gpu_hop(s, fla, prec, srcs, eo_out=ODD):
    n = len(srcs)
    vin = [s.vec(x) for x in srcs]
    vout = [s.zeros() for _ in range(n)]
    assert s.lib.QOPX_D3_asqtad_hop(fla, prec, n, (P * n)(*vout), (P * n)(*vin), eo_out) == 0
    out = [s.get(v) for v in vout]
    for v in vin + vout:
        s.free(v)
    return out


class TestConfig2:
    """Both tests share one 32^3x64 session, closed before the tests below
    open sessions of their own (QOP_init holds one lattice at a time)."""

    def test_hop_fp32_fp16_against_oracle_32x32x32x64(self, c2, oracle):
        s, fla, h = c2["s"], c2["fla"], c2["h"]
        V = s.V
        srcs = [q.gen_source(L2, 2000 + i, EVEN) for i in range(16)]
        refs = [oracle.dslash(h, 0.0, x, ODD, EVEN)[V // 2:] for x in srcs[:4]]
        errs = {}
        for prec, bound in ((32, 1e-6), (16, 1e-3)):
            one = gpu_hop(s, fla, prec, srcs[:1])[0][V // 2:]
            many = gpu_hop(s, fla, prec, srcs)
            e = [rel(many[i][V // 2:], refs[i]) for i in range(4)]
            errs[prec] = max(e)
            assert max(e) <= bound, (prec, e)
            # the single-RHS launch gives the same bits as slot 0 of the 16-RHS one
            assert np.array_equal(one, many[0][V // 2:])
        # each format's error is its own: FP16 well above FP32 rounding
        assert errs[16] > 30 * errs[32]
        print(f"hop relative L2 vs FP64 oracle at 32^3x64: FP32 {errs[32]:.2e}, FP16 {errs[16]:.2e}")


    def test_fp16_batch_solves_to_1e10_32x32x32x64(self, c2, oracle):
        """The headline workload: 16 sources, FP16 sloppy batched CG with FP64
        reliable updates, m = 0.01, every solve to FP64 true residual 1e-10 by
        the oracle's operator."""
        from test_gpu_invert import batch_api
        s, fla, h = c2["s"], c2["fla"], c2["h"]
        m, tol = 0.01, 1e-10
        srcs = [q.gen_source(L2, 2000 + i, EVEN) for i in range(16)]
        ok, outs, ras = batch_api(s, fla, srcs, [m] * 16, tol, 16, capacity=16)
        assert ok == 16
        worst = 0.0
        for b, o, r in zip(srcs, outs, ras):
            tr = oracle.true_resid_even(h, m, o, b)
            assert tr <= tol, tr
            assert r.final_rsq == pytest.approx(tr * tr, rel=1e-2)
            worst = max(worst, tr)
        print(f"32^3x64 x16 FP16 batch: worst FP64 true residual (oracle) {worst:.3e}")


MASSES_C3 = [0.005, 0.00846907, 0.014345, 0.0242978, 0.041156, 0.0697106, 0.118077, 0.2]


def test_multi_mass_batch_16x8_24x24x24x48(oracle):
    from test_gpu_invert import batch_api
    L = [24, 24, 24, 48]
    U = q.gen_gauge(L, 1000, 0.3)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    del U
    h = oracle.links(L, fat, lng)
    s = q.Session(L)
    try:
        fla = s.links_from_raw(fat, lng)
        del fat, lng
        srcs = [q.gen_source(L, 2300 + i, EVEN) for i in range(16)]
        tol = 1e-10
        ok, outs, ras = batch_api(s, fla, srcs, MASSES_C3, tol, 16, multi=True, capacity=16)
        assert ok == 16 * len(MASSES_C3)
        worst = 0.0
        for i, b in enumerate(srcs):
            for j, m in enumerate(MASSES_C3):
                k = i * len(MASSES_C3) + j
                tr = oracle.true_resid_even(h, m, outs[k], b)
                assert tr <= tol, (i, m, tr)
                assert ras[k].final_rsq == pytest.approx(tr * tr, rel=1e-2)
                worst = max(worst, tr)
        print(f"24^3x48 16x8 multi-mass: worst FP64 true residual (oracle) {worst:.3e}")
        s.lib.QOP_D3_asqtad_destroy_L(fla)
    finally:
        oracle.free_links(h)
        s.close()


MASSES_C5 = [0.002, 0.00330388, 0.00545782, 0.00901599, 0.0148939, 0.0246038, 0.0406441,
             0.0671416, 0.110914, 0.183223, 0.302674, 0.5]


def test_hisq_12_masses_48x48x48x96(oracle):
    from test_gpu_hisq import gpu_hisq
    from test_gpu_invert import batch_api
    L = [48, 48, 48, 96]
    U = q.gen_gauge(L, 1000, 0.3)
    s = q.Session(L)
    h = None
    try:
        hl, info = gpu_hisq(s, U, q.hisq_coeffs())
        del U
        gc.collect()
        assert hl and info.status == 0
        fla = s.lib.QOP_D3_get_asqtad_links_from_hisq(hl)[0]
        src = q.gen_source(L, 2500, EVEN)
        tol = 1e-10
        ok, outs, ras = batch_api(s, fla, [src], MASSES_C5, tol, 16, multi=True, capacity=12)
        assert ok == len(MASSES_C5)
        fat, lng = s.extract_links(fla)
        s.lib.QOP_D3_hisq_destroy_L(hl)
        h = oracle.links(L, fat, lng)
        del fat, lng
        gc.collect()
        worst = 0.0
        for j, m in enumerate(MASSES_C5):
            tr = oracle.true_resid_even(h, m, outs[j], src)
            assert tr <= tol, (m, tr)
            worst = max(worst, tr)
        print(f"48^3x96 HISQ x12 masses: worst FP64 true residual (oracle) {worst:.3e}")
    finally:
        if h:
            oracle.free_links(h)
        s.close()


def test_serving_mix_three_lattices(oracle):
    """Requests over 16^4, 24^3x48 and 32^3x64 (one resident configuration
    each), EVEN and ODD, masses log-uniform on [0.005, 0.1], tolerances 1e-8 /
    1e-10: alone and in the concurrent mix the same bits; each reaches its
    tolerance by the oracle's operator."""
    lats = [[16, 16, 16, 16], [24, 24, 24, 48], L2]
    srv = q.Server(capacity=4, prec=16)
    hs, cfg_of = [], []
    try:
        for li, L in enumerate(lats):
            U = q.gen_gauge(L, 1000 + li, 0.3)
            cfg_of.append(srv.add_gauge(L, U, MILC))
            oracle.rephase_milc(L, U)
            fat, lng = oracle.make_imp_links(L, U, MILC)
            hs.append(oracle.links(L, fat, lng))
            del U, fat, lng
        rng = np.random.default_rng(31337)
        reqs = []
        for i in range(6):
            li = i % 3
            eo = int(rng.integers(2))
            m = float(np.exp(rng.uniform(np.log(0.005), np.log(0.1))))
            tol = [1e-8, 1e-10][int(rng.integers(2))]
            reqs.append(dict(li=li, eo=eo, m=m, tol=tol, src=q.gen_source(lats[li], 2600 + i, eo)))
        alone = []
        for r in reqs:
            rid = srv.submit(cfg_of[r["li"]], r["src"], r["m"], r["tol"], r["eo"])
            out, st, res, lat = srv.wait(rid)
            assert st == 0
            alone.append(out)
        ids = [srv.submit(cfg_of[r["li"]], r["src"], r["m"], r["tol"], r["eo"]) for r in reqs]
        for r, rid, a in zip(reqs, ids, alone):
            out, st, res, lat = srv.wait(rid)
            assert st == 0
            assert np.array_equal(out, a)
            tr = oracle.true_resid_even(hs[r["li"]], r["m"], out, r["src"], r["eo"])
            assert tr <= r["tol"], (r, tr)
    finally:
        srv.close()
        for h in hs:
            oracle.free_links(h)
