"""The batched solver's result does not depend on the multi-RHS hop launch
shape (RHS per thread, block size, XCD remap, non-temporal loads and stores,
block order, x0-tiled site order, LDS staging of the block's source rows or
of its x0 planes in the plane-blocked order, buffer loads or the plain-load
kernel that lattices beyond the buffer range take):
each forced shape gives bit-identical solutions and iteration counts, all at FP64 true residual 1e-10 (oracle operator).
A forced shape also applies to the FP64 reliable-update hops over the
compacted slot list wherever it is valid for them (2 or 4 RHS per thread
with LDS staging, up to 8 without), so their staged and unstaged forms are
covered too.  The
second lattice has x3 rows of 5 sites, so staged blocks end inside a row and
gathers switch between the LDS tile and memory within a wavefront."""
import ctypes as C

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC

pytestmark = pytest.mark.gpu
EVEN = 0

# (RHS per thread, block, XCD remap, non-temporal, RHS-innermost, x0 tile,
#  LDS staging, plain loads)
SHAPES = [
    (4, 128, 1, 0, 1, 0, 0, 0),
    (4, 128, 1, 0, 1, 512, 0, 0),
    (2, 64, 0, 1, 0, 1024, 0, 0),
    (1, 256, 1, 0, 0, 512, 0, 0),
    (8, 64, 1, 1, 1, 1024, 0, 0),
    (4, 128, 1, 0, 1, 0, 1, 0),
    (2, 64, 0, 0, 0, 0, 1, 0),
    (8, 256, 1, 0, 1, 0, 1, 0),
    (1, 64, 1, 0, 1, 0, 1, 0),
    (4, 128, 1, 0, 1, 512, 1, 0),
    (4, 256, 1, 0, 1, 0, 2, 0),
    (2, 128, 0, 0, 0, 0, 2, 0),
    (4, 512, 1, 0, 1, 0, 2, 0),
    (8, 256, 1, 0, 1, 0, 2, 0),
    (1, 512, 0, 0, 1, 0, 2, 0),
    (2, 256, 0, 0, 0, 1024, 1, 0),
    (1, 256, 1, 0, 1, 0, 0, 1),
    (1, 64, 0, 0, 0, 512, 0, 1),
]


# x0 slab of 2048 sites (tiles of 512 and 1024 apply); x3 rows of 5 sites
@pytest.mark.parametrize("L", [[8, 16, 16, 16], [6, 6, 6, 10]])
def test_batched_solves_independent_of_launch_shape(oracle, L):
    from test_gpu_invert import batch_api
    U = q.gen_gauge(L, 5001, 0.3)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    h = oracle.links(L, fat, lng)
    s = q.Session(L)
    try:
        fla = s.links_from_raw(fat, lng)
        srcs = [q.gen_source(L, 5100 + i, EVEN) for i in range(4)]
        masses = [0.01, 0.05, 0.2, 0.02]
        ref = None
        for prec in (16, 32):
            ref = None
            for shape in SHAPES:
                s.lib.QOPX_hop_shape_override((C.c_int * 8)(*shape))
                ok, outs, ras = batch_api(s, fla, srcs, masses, 1e-10, prec, capacity=4)
                assert ok == 4, shape
                its = [r.final_iter for r in ras]
                if ref is None:
                    ref = (outs, its)
                    for b, m, o in zip(srcs, masses, outs):
                        assert oracle.true_resid_even(h, m, o, b) <= 1e-10
                else:
                    assert its == ref[1], (prec, shape)
                    for o, r in zip(outs, ref[0]):
                        assert np.array_equal(o, r), (prec, shape)
        s.lib.QOP_D3_asqtad_destroy_L(fla)
    finally:
        s.lib.QOPX_hop_shape_override(None)
        oracle.free_links(h)
        s.close()
