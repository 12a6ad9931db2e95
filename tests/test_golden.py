"""Golden fixtures (tests/golden/asqtad_4x4x4x8.npz, written by
tests/golden/make_golden.py from the oracle): the seeded inputs must
regenerate bit for bit, the oracle must reproduce its frozen outputs, and the
GPU library -- taking the same raw gauge field through rephase_G, fattening,
Dslash, CG and multi-mass CG on the device -- must match them.

Tolerances: checksums and Dslash relative 1e-13 (FP64, different summation
order), solutions relative 1e-8 (two solves to |r|/|b| = 1e-12 of a system with
condition number ~1e3), iteration counts within 2 (oracle reductions depend
on the OpenMP thread count)."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC

EVEN, ODD, EVENODD = 0, 1, 2
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "asqtad_4x4x4x8.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def sha(arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def inputs(gold):
    L = [int(x) for x in gold["L"]]
    U = q.gen_gauge(L, int(gold["gauge_seed"]), float(gold["eps"]))
    srcs = [q.gen_source(L, int(s), EVENODD) for s in gold["src_seeds"]]
    return L, U, srcs


def test_inputs_regenerate(gold, oracle):
    L, U, srcs = inputs(gold)
    assert sha(U) == str(gold["gauge_sha256"])
    assert [sha([s]) for s in srcs] == [str(x) for x in gold["src_sha256"]]
    assert abs(oracle.plaquette(L, U) - float(gold["plaquette"])) < 1e-14
    oracle.rephase_milc(L, U)
    assert sha(U) == str(gold["rephased_sha256"])


def test_oracle_reproduces_links_and_dslash(gold, oracle):
    L, U, srcs = inputs(gold)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    sites = gold["sample_sites"]
    for mu in range(4):
        assert np.allclose([fat[mu][..., 0].sum(), fat[mu][..., 1].sum()], gold["fat_sum"][mu],
                           rtol=1e-12, atol=1e-12)
        assert abs(np.sqrt((fat[mu] ** 2).sum()) - gold["fat_norm"][mu]) <= 1e-13 * gold["fat_norm"][mu]
        assert abs(np.sqrt((lng[mu] ** 2).sum()) - gold["long_norm"][mu]) <= 1e-13 * gold["long_norm"][mu]
        assert rel(fat[mu][sites], gold["fat_sites"][mu]) <= 1e-13
        assert rel(lng[mu][sites], gold["long_sites"][mu]) <= 1e-13
    h = oracle.links(L, fat, lng)
    try:
        for i, s in enumerate(srcs):
            d = oracle.dslash(h, float(gold["dslash_mass"]), s, EVENODD, EVENODD)
            assert rel(d, gold["dslash"][i]) <= 1e-13
    finally:
        oracle.free_links(h)


def test_oracle_reproduces_solves(gold, oracle):
    L, U, srcs = inputs(gold)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    h = oracle.links(L, fat, lng)
    try:
        b = srcs[0].copy()
        b[b.shape[0] // 2:] = 0
        x, ra, _ = oracle.invert(h, float(gold["solve_mass"]), b, EVEN, float(gold["rsqmin"]))
        assert rel(x, gold["solve_x"]) <= 1e-8
        assert abs(ra.final_iter - int(gold["solve_iters"])) <= 2
        outs, ras, _ = oracle.invert_multi(h, [list(gold["multi_masses"])], [b], EVEN,
                                           float(gold["rsqmin"]))
        for k in range(len(outs)):
            assert rel(outs[k], gold["multi_x"][k]) <= 1e-8
    finally:
        oracle.free_links(h)


@pytest.mark.gpu
def test_gpu_matches_golden(gold):
    """Raw gauge field -> GPU rephase_G -> GPU fattening -> Dslash / CG /
    multi-mass CG, all on the device, against the frozen oracle outputs."""
    from test_gpu_invert import invert_multi, solve
    L, U, srcs = inputs(gold)
    s = q.Session(L)
    try:
        g = s.gauge(U)
        r0, bc, sign, keep = q.milc_rephase_args(L)
        s.lib.QOP_D3_rephase_G(g, r0, C.byref(bc), C.byref(sign))
        fla, _ = s.links_from_gauge(g)
        fat, lng = s.extract_links(fla)
        for mu in range(4):
            assert rel(fat[mu][gold["sample_sites"]], gold["fat_sites"][mu]) <= 1e-13
            assert abs(np.sqrt((lng[mu] ** 2).sum()) - gold["long_norm"][mu]) <= 1e-13 * gold["long_norm"][mu]
        for i, src in enumerate(srcs):
            vin, vout = s.vec(src), s.zeros()
            s.lib.QOP_D3_asqtad_dslash(C.byref(q.QOPInfo()), fla, float(gold["dslash_mass"]), vout,
                                       vin, EVENODD, EVENODD)
            assert rel(s.get(vout), gold["dslash"][i]) <= 1e-13
            s.free(vin)
            s.free(vout)
        b = srcs[0].copy()
        b[b.shape[0] // 2:] = 0
        x, ra, _ = solve(s, fla, float(gold["solve_mass"]), b, EVEN, float(gold["rsqmin"]))
        assert rel(x, gold["solve_x"]) <= 1e-8
        assert abs(ra.final_iter - int(gold["solve_iters"])) <= 2
        res, _ = invert_multi(s, fla, [list(gold["multi_masses"])], [b], EVEN, float(gold["rsqmin"]))
        for k in range(len(res[0])):
            assert rel(res[0][k], gold["multi_x"][k]) <= 1e-8
        s.lib.QOP_D3_asqtad_destroy_L(fla)
        s.lib.QOP_D3_destroy_G(g)
    finally:
        s.close()
