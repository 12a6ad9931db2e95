"""The force fixtures (tests/golden/force_4x4x4x8.npz, written by
tests/golden/make_golden_force.py from the finite-difference-pinned oracle):
the seeded inputs regenerate, the oracle reproduces the asqtad and HISQ
forces (per-direction sums and norms, full matrices at sample sites) to
1e-12, and the GPU library reproduces them to 1e-10 through the C ABI."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden_force as mg  # noqa: E402

GOLD = os.path.join(HERE, "golden", "force_4x4x4x8.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD, allow_pickle=False))


def close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max()) <= tol * max(float(np.abs(b).max()), 1e-300)


def test_inputs_regenerate(gold):
    U, xs = mg.inputs()
    assert mg.sha(U + xs) == str(gold["inputs_sha256"])


def test_oracle_reproduces_forces(gold):
    out = mg.generate()
    for k in ("asqtad", "hisq"):
        for part in ("sum", "norm", "sample"):
            assert close(out[f"{k}_{part}"], gold[f"{k}_{part}"], 1e-12), (k, part)
    assert int(out["hisq_filtered"]) == int(gold["hisq_filtered"])


@pytest.mark.gpu
def test_gpu_matches_force_fixtures(gold, oracle):
    import qopqdp_amd as q
    L = mg.L
    V = int(np.prod(L))
    U, xs = mg.inputs()
    oracle.rephase_milc(L, U)
    s = q.Session(L)
    try:
        g = s.gauge(U)
        vs = [s.vec(x) for x in xs]
        zero = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        f = s.force(zero)
        info = s.asqtad_force(g, f, mg.EPS_TERMS, vs, mg.MILC)
        assert info.status == 0
        sm, nm, sp = mg.summary(s.get_force(f))
        assert close(sm, gold["asqtad_sum"], 1e-10) and close(nm, gold["asqtad_norm"], 1e-10)
        assert close(sp, gold["asqtad_sample"], 1e-10)
        s.lib.QOP_D3_destroy_F(f)
        hc = q.hisq_coeffs(eps_naik=tuple(mg.EPS_NAIK))
        hinfo = q.QOPInfo()
        flh = s.lib.QOP_D3_hisq_create_L_from_G(C.byref(hinfo), C.byref(hc), g)
        assert hinfo.status == 0
        f = s.force(zero)
        n = len(vs)
        info = q.QOPInfo()
        s.lib.QOP_D3_hisq_force_multi(C.byref(info), flh, f, C.byref(hc), (C.c_double * n)(*mg.EPS_TERMS),
                                      (C.c_void_p * n)(*vs), (C.c_int * len(mg.GROUPS))(*mg.GROUPS))
        assert info.status == 0
        assert info.count2 == int(gold["hisq_filtered"])
        sm, nm, sp = mg.summary(s.get_force(f))
        assert close(sm, gold["hisq_sum"], 1e-10) and close(nm, gold["hisq_norm"], 1e-10)
        assert close(sp, gold["hisq_sample"], 1e-10)
        s.lib.QOP_D3_destroy_F(f)
        s.lib.QOP_D3_hisq_destroy_L(flh)
        for v in vs:
            s.free(v)
        s.lib.QOP_D3_destroy_G(g)
    finally:
        s.close()
