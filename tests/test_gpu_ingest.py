"""API-edge transfers from page-locked host buffers (QOPX_host_alloc): a raw
vector round trip through QOP_D3_create_V_from_raw / extract_V_to_raw is
exact, as from ordinary (pageable) memory."""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q

pytestmark = pytest.mark.gpu


def pinned_array(lib, shape):
    n = int(np.prod(shape))
    p = lib.QOPX_host_alloc(n * 8)
    assert p
    a = np.ctypeslib.as_array((C.c_double * n).from_address(p)).reshape(shape)
    return p, a


@pytest.mark.parametrize("eo", [q.QOP_EVEN, q.QOP_ODD, q.QOP_EVENODD])
def test_pinned_raw_vector_round_trip(eo):
    L = [8, 8, 8, 16]
    s = q.Session(L)
    try:
        lib = s.lib
        src = q.gen_source(L, 31, q.QOP_EVENODD)
        p_in, a_in = pinned_array(lib, src.shape)
        p_out, a_out = pinned_array(lib, src.shape)
        a_in[...] = src
        a_out[...] = -1.0
        v = lib.QOP_D3_create_V_from_raw(a_in, eo)  # views of the pinned buffers
        lib.QOP_D3_extract_V_to_raw(a_out, v, eo)
        Vh = s.V // 2
        sl = {q.QOP_EVEN: slice(0, Vh), q.QOP_ODD: slice(Vh, 2 * Vh), q.QOP_EVENODD: slice(0, 2 * Vh)}[eo]
        assert np.array_equal(a_out[sl], src[sl])
        if eo != q.QOP_EVENODD:  # sites outside the subset are not written
            other = slice(Vh, 2 * Vh) if eo == q.QOP_EVEN else slice(0, Vh)
            assert (a_out[other] == -1.0).all()
        lib.QOP_D3_destroy_V(v)
        lib.QOPX_host_free(p_in)
        lib.QOPX_host_free(p_out)
    finally:
        s.close()
