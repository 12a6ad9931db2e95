"""GPU parity of the asqtad fermion force (QOP_D3_/QOP_F3_asqtad_force[_multi],
lib/asqtad_force_p.c:41-63, fnmat form lib/asqtad_force_fnmat_p.c:474-546)
through the C ABI.

* against the oracle (oracle/force_oracle.c, itself pinned by finite
  differences of the smeared action) on small lattices, for the MILC
  coefficients and for coefficient sets that take each branch of the
  smearing derivative: relative L2 <= 1e-12 (FP64), <= 1e-5 (FP32 API: FP32
  inputs and output, FP64 arithmetic);
* on the GPU alone at 8^3x16: the force is the derivative of
  S(U) = sum_i eps_i |D_oe(U) X_i|^2 with the links smeared and the Dslash
  applied on the GPU (central differences, relative 1e-6);
* at 32^3x64 (the size of the solver workload): traceless antihermitian to
  1e-12 and exactly linear in eps (force(2 eps) == 2 force(eps) bit for bit).
"""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC, from_raw, gauge_numpy, to_raw

pytestmark = pytest.mark.gpu
EVEN, ODD, EVENODD = 0, 1, 2

ZERO = dict(one_link=0.0, three_staple=0.0, five_staple=0.0, seven_staple=0.0, lepage=0.0, naik=0.0)
SETS = {
    "milc": MILC,
    "one_link": dict(ZERO, one_link=0.7),
    "three_staple": dict(ZERO, one_link=0.2, three_staple=0.3),
    "five_staple": dict(ZERO, five_staple=0.2),
    "seven_only": dict(ZERO, seven_staple=0.2),
    "lepage_only": dict(ZERO, lepage=0.25),
    "five_lepage": dict(ZERO, five_staple=0.1, lepage=-0.05),
    "naik_only": dict(ZERO, naik=0.4),
}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vectors(L, seed, n):
    V = int(np.prod(L))
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((V, 3, 2)) for _ in range(n)]


def gpu_force(s, U, F0, coeffs, eps, xs, prec="D"):
    g = s.gauge(U, prec)
    f = s.force(F0, prec)
    vs = [s.vec(x, EVENODD, prec) for x in xs]
    try:
        info = s.asqtad_force(g, f, eps, vs, coeffs, prec)
        assert info.status == 0
        return s.get_force(f, prec), info
    finally:
        for v in vs:
            s.free(v, prec)
        (s.lib.QOP_D3_destroy_F if prec == "D" else s.lib.QOP_F3_destroy_F)(f)
        (s.lib.QOP_D3_destroy_G if prec == "D" else s.lib.QOP_F3_destroy_G)(g)


@pytest.fixture
def small():
    # one QOP session at a time: the tests below open other lattices
    L = [4, 4, 6, 8]
    s = q.Session(L)
    yield L, s
    s.close()


@pytest.mark.parametrize("name", list(SETS))
def test_force_matches_oracle(oracle, small, name):
    L, s = small
    coeffs = SETS[name]
    U = gauge_numpy(L, 8100, 0.5)
    xs = vectors(L, 8200, 3)
    eps = [0.7, -0.4, 1.3]
    rng = np.random.default_rng(8300)
    F0 = [rng.standard_normal((int(np.prod(L)), 3, 3, 2)) * 0.1 for _ in range(4)]
    ref = oracle.asqtad_force(L, U, coeffs, eps, xs, force=[f.copy() for f in F0])
    got, info = gpu_force(s, U, F0, coeffs, eps, xs)
    for mu in range(4):
        d = rel(got[mu] - F0[mu], ref[mu] - F0[mu])
        assert d <= 1e-12, (name, mu, d)
    assert info.final_flop > 0 and info.final_sec > 0


def test_force_single_source_entry_point(small):
    L, s = small
    U = gauge_numpy(L, 8400, 0.4)
    x = vectors(L, 8500, 1)[0]
    F0 = [np.zeros((int(np.prod(L)), 3, 3, 2)) for _ in range(4)]
    multi, _ = gpu_force(s, U, F0, MILC, [0.9], [x])
    g = s.gauge(U)
    f = s.force(F0)
    v = s.vec(x)
    info = q.QOPInfo()
    c = q.QOPAsqtadCoeffs(**MILC)
    s.lib.QOP_D3_asqtad_force(C.byref(info), g, f, C.byref(c), 0.9, v)
    single = s.get_force(f)
    s.free(v)
    s.lib.QOP_D3_destroy_F(f)
    s.lib.QOP_D3_destroy_G(g)
    for mu in range(4):
        assert np.array_equal(single[mu], multi[mu])


def test_force_fp32_api(oracle, small):
    L, s = small
    U = [u.astype(np.float32) for u in gauge_numpy(L, 8600, 0.4)]
    xs = [x.astype(np.float32) for x in vectors(L, 8700, 2)]
    eps = [0.5, 0.25]
    F0 = [np.zeros((int(np.prod(L)), 3, 3, 2), dtype=np.float32) for _ in range(4)]
    ref = oracle.asqtad_force(L, [u.astype(np.float64) for u in U], MILC, eps,
                              [x.astype(np.float64) for x in xs])
    got, _ = gpu_force(s, U, F0, MILC, eps, xs, prec="F")
    for mu in range(4):
        assert got[mu].dtype == np.float32
        assert rel(got[mu], ref[mu]) <= 1e-5


def test_force_set_opts(small):
    _, s = small

    def set_opts(**kw):
        opts = (q.QOPOpt * len(kw))()
        keep = []
        for i, (k, v) in enumerate(kw.items()):
            keep.append(k.encode())
            opts[i] = q.QOPOpt(keep[-1], float(v), None)
        return s.lib.QOP_asqtad_force_set_opts(opts, len(kw))

    assert set_opts(fnmat_src_min=1, veclength=8) == 0
    assert set_opts(fnmat_src_min=-1) != 0
    assert set_opts(veclength=0) != 0
    assert set_opts(fnmat_src_min=2, veclength=4) == 0


def test_force_is_the_derivative_on_the_gpu():
    """S(U) = sum_i eps_i |D_oe X_i|^2 entirely on the GPU: the links are
    smeared from U by QOP_D3_asqtad_create_L_from_G and D_oe applied by
    QOP_D3_asqtad_dslash; x_i = X_i (even) + D_oe X_i (odd)."""
    L = [8, 8, 8, 16]
    V = int(np.prod(L))
    s = q.Session(L)
    try:
        U = gauge_numpy(L, 8800, 0.5)
        rng = np.random.default_rng(8900)
        Xs = []
        for _ in range(2):
            X = rng.standard_normal((V, 3, 2))
            X[V // 2:] = 0
            Xs.append(X)
        eps = [0.6, -0.3]

        def d_oe(Uc):
            g = s.gauge(Uc)
            fla, info = s.links_from_gauge(g)
            assert info.status == 0
            out = []
            for X in Xs:
                vin = s.vec(X)
                vout = s.zeros()
                inf = q.QOPInfo()
                s.lib.QOP_D3_asqtad_dslash(C.byref(inf), fla, 0.0, vout, vin, ODD, EVEN)
                out.append(s.get(vout))
                s.free(vin)
                s.free(vout)
            s.lib.QOP_D3_asqtad_destroy_L(fla)
            s.lib.QOP_D3_destroy_G(g)
            return out

        def action(Uc):
            return sum(e * float(np.sum(Y[V // 2:] ** 2)) for e, Y in zip(eps, d_oe(Uc)))

        xs = []
        for X, Y in zip(Xs, d_oe(U)):
            x = X.copy()
            x[V // 2:] = Y[V // 2:]
            xs.append(x)
        F0 = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        F, _ = gpu_force(s, U, F0, MILC, eps, xs)
        h = 1e-4
        s0 = abs(action(U))  # rounding of S enters fd as ~1e-16 |S| / h
        for k in range(4):
            z = int(rng.integers(V))
            mu = int(rng.integers(4))
            gm = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
            T = 0.5 * (gm + np.conj(gm.T))
            T -= np.trace(T) / 3 * np.eye(3)
            w, vv = np.linalg.eigh(T)
            u0 = from_raw(U[mu][z])
            sv = []
            for t in (h, -h):
                Up = [u.copy() for u in U]
                Up[mu][z] = to_raw((((vv * np.exp(1j * t * w)) @ np.conj(vv.T)) @ u0)[None])[0]
                sv.append(action(Up))
            fd = (sv[0] - sv[1]) / (2 * h)
            pred = 0.5 * np.imag(np.trace(T @ from_raw(F[mu][z])))
            assert abs(fd - pred) <= 1e-6 * abs(pred) + 1e-15 * s0 / h, (z, mu, fd, pred)
    finally:
        s.close()


def test_force_full_size_properties():
    """32^3x64: traceless antihermitian, exactly linear in eps."""
    L = [32, 32, 32, 64]
    V = int(np.prod(L))
    s = q.Session(L)
    try:
        U = q.gen_gauge(L, 9000, 0.3)  # held until the upload has read it
        g = s.gauge(U)
        x = s.vec(q.gen_source(L, 9100, EVENODD))
        F0 = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        out = []
        for e in (1.0, 2.0):
            f = s.force(F0)
            info = s.asqtad_force(g, f, [e], [x])
            assert info.status == 0
            out.append(s.get_force(f))
            s.lib.QOP_D3_destroy_F(f)
        for mu in range(4):
            assert np.array_equal(out[1][mu], 2 * out[0][mu])
            m = from_raw(out[0][mu][::97])
            scale = np.abs(m).max()
            assert scale > 0
            assert np.abs(m + np.conj(np.transpose(m, (0, 2, 1)))).max() <= 1e-12 * scale
            assert np.abs(np.trace(m, axis1=1, axis2=2)).max() <= 1e-12 * scale
        s.free(x)
        s.lib.QOP_D3_destroy_G(g)
    finally:
        s.close()
