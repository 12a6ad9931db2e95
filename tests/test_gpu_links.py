"""GPU link preparation through the C ABI: rephase_G, asqtad fattening
(create_L_from_G), and the free-field spectrum of the GPU-built operator.
Each test runs its own QOP session (one lattice per QOP_init)."""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC

pytestmark = pytest.mark.gpu
EVEN, ODD, EVENODD = 0, 1, 2


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_rephase_and_fattening_on_gpu(oracle):
    L = [4, 4, 4, 6]
    U = q.gen_gauge(L, 1002, 0.3)
    s = q.Session(L)
    try:
        g = s.gauge(U)
        r0, bc, sign, keep = q.milc_rephase_args(L)
        s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
        Ug = [np.zeros_like(u) for u in U]
        s.lib.QOP_D3_extract_G_to_raw(q.ptr_array(Ug), g, EVENODD)
        Ur = [u.copy() for u in U]
        oracle.rephase_milc(L, Ur)
        for mu in range(4):
            assert np.array_equal(Ug[mu], Ur[mu])
        fla, info = s.links_from_gauge(g)
        assert info.final_flop == 61632 * s.V
        fat, lng = s.extract_links(fla)
        ofat, olng = oracle.make_imp_links(L, Ur, MILC)
        for mu in range(4):
            assert rel(fat[mu], ofat[mu]) <= 1e-13
            assert rel(lng[mu], olng[mu]) <= 1e-13
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        s.lib.QOP_D3_destroy_G(g)
    finally:
        s.close()


def test_free_field_spectrum_gpu(oracle):
    """A psi = (m^2 + sum s_mu^2) psi on even plane waves, U = 1 (GPU fattening)."""
    L = [4, 4, 4, 8]
    V = int(np.prod(L))
    eye = np.zeros((V, 3, 3, 2))
    eye[:, [0, 1, 2], [0, 1, 2], 0] = 1
    s = q.Session(L)
    try:
        g = s.gauge([eye] * 4)
        r0, bc, sign, keep = q.milc_rephase_args(L)
        s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
        fla, _ = s.links_from_gauge(g)
        x = q.raw_coords_all(L)
        p = np.array([2 * np.pi * 1 / 4, 0, 2 * np.pi * 3 / 4, 5 * np.pi / 8])
        ph = np.exp(1j * (x @ p))
        v = np.zeros((V, 3, 2))
        v[: V // 2, 1, 0] = ph.real[: V // 2]
        v[: V // 2, 1, 1] = ph.imag[: V // 2]
        m = 0.05
        sp = 9 / 8 * np.sin(p) - 1 / 24 * np.sin(3 * p)
        lam = m * m + np.sum(sp * sp)
        vin, tmp, out = s.vec(v), s.zeros(), s.zeros()
        info = q.QOPInfo()
        s.lib.QOP_D3_asqtad_dslash(info, fla, m, tmp, vin, ODD, EVEN)
        s.lib.QOP_D3_asqtad_dslash(info, fla, m, out, tmp, EVEN, ODD)
        Av = m * m * v - s.get(out)
        np.testing.assert_allclose(Av[: V // 2], lam * v[: V // 2], atol=1e-13)
        for h_ in (vin, tmp, out):
            s.free(h_)
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        s.lib.QOP_D3_destroy_G(g)
    finally:
        s.close()
