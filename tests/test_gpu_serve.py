"""The QOPX_serve_* request server on the GPU: requests over two lattices
with two resident configurations each, EVEN and ODD subsets, mixed masses and
tolerances.  Each request must give bit-identical output alone and in the
concurrent mix (per-slot arithmetic, fixed-order reductions), reach its FP64
true relative residual (checked with the oracle's operator built from the same
gauge field), and the same must hold when the HBM budget forces batches to be
evicted and re-created, and for the FP64 one-at-a-time server."""
import numpy as np
import pytest

import qopqdp_amd as q
from conftest import MILC

pytestmark = pytest.mark.gpu
EVEN, ODD = 0, 1

LATTICES = [[4, 4, 4, 8], [4, 4, 6, 8]]


@pytest.fixture(scope="module")
def world(oracle):
    cfgs = []
    for li, L in enumerate(LATTICES):
        for k in range(2):
            U = q.gen_gauge(L, 3000 + 10 * li + k, 0.3)
            Ur = [u.copy() for u in U]
            oracle.rephase_milc(L, Ur)
            fat, lng = oracle.make_imp_links(L, Ur, MILC)
            cfgs.append(dict(L=L, U=U, h=oracle.links(L, fat, lng)))
    rng = np.random.default_rng(31337)
    reqs = []
    for i in range(12):
        c = int(rng.integers(len(cfgs)))
        eo = int(rng.integers(2))
        m = float(np.exp(rng.uniform(np.log(0.02), np.log(0.2))))
        tol = [1e-8, 1e-10][int(rng.integers(2))]
        reqs.append(dict(cfg=c, eo=eo, mass=m, tol=tol,
                         src=q.gen_source(cfgs[c]["L"], 3100 + i, eo)))
    yield cfgs, reqs
    for c in cfgs:
        oracle.free_links(c["h"])


def make_server(cfgs, **kw):
    srv = q.Server(**kw)
    for c in cfgs:
        srv.add_gauge(c["L"], c["U"], MILC)
    return srv


def check_resid(oracle, cfgs, r, out):
    tr = oracle.true_resid_even(cfgs[r["cfg"]]["h"], r["mass"], out, r["src"], r["eo"])
    assert tr <= r["tol"], (r["cfg"], r["eo"], r["mass"], r["tol"], tr)
    return tr


def test_alone_equals_mix(world, oracle):
    cfgs, reqs = world
    srv = make_server(cfgs, capacity=4, prec=16)
    try:
        alone = []
        for r in reqs:
            rid = srv.submit(r["cfg"], r["src"], r["mass"], r["tol"], r["eo"])
            out, st, res, lat = srv.wait(rid)
            assert st == 0
            check_resid(oracle, cfgs, r, out)
            assert res.final_rsq <= r["tol"] ** 2
            alone.append((out, res.final_iter))
        ids = [srv.submit(r["cfg"], r["src"], r["mass"], r["tol"], r["eo"]) for r in reqs]
        for r, rid, (a, it) in zip(reqs, ids, alone):
            out, st, res, lat = srv.wait(rid)
            assert st == 0
            assert np.array_equal(out, a)
            assert res.final_iter == it
            # sites outside the request's subset are not written
            V = out.shape[0]
            other = slice(V // 2, V) if r["eo"] == EVEN else slice(0, V // 2)
            assert not out[other].any()
        st = srv.stats()
        assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    finally:
        srv.close()


def test_tight_budget_evicts_and_matches(world, oracle):
    """A budget with room for the links and about one batch: batches of
    different configurations take turns, results stay bit-identical."""
    cfgs, reqs = world
    probe = make_server(cfgs, capacity=4, prec=16)
    used = [probe.stats()["bytes_used"]]
    for li in range(len(LATTICES)):  # one batch per lattice size
        r = next(r for r in reqs if cfgs[r["cfg"]]["L"] == LATTICES[li])
        probe.wait(probe.submit(r["cfg"], r["src"], r["mass"], r["tol"], r["eo"]))
        used.append(probe.stats()["bytes_used"])
    rid = probe.submit(reqs[0]["cfg"], reqs[0]["src"], reqs[0]["mass"], reqs[0]["tol"], reqs[0]["eo"])
    ref0 = probe.wait(rid)[0]
    probe.close()
    batch = max(np.diff(used))
    srv = make_server(cfgs, capacity=4, prec=16, max_bytes=int(used[0] + 1.5 * batch))
    try:
        ids = [srv.submit(r["cfg"], r["src"], r["mass"], r["tol"], r["eo"]) for r in reqs]
        outs = [srv.wait(i) for i in ids]
        for r, (out, st, res, lat) in zip(reqs, outs):
            assert st == 0
            check_resid(oracle, cfgs, r, out)
        assert np.array_equal(outs[0][0], ref0)
        assert srv.stats()["batches"] <= 1
    finally:
        srv.close()


def test_fp64_server_and_zero_source(world, oracle):
    cfgs, reqs = world
    srv = make_server(cfgs, capacity=1, prec=64)
    try:
        for r in reqs[:4]:
            out, st, res, lat = srv.wait(srv.submit(r["cfg"], r["src"], r["mass"], r["tol"], r["eo"]))
            assert st == 0
            check_resid(oracle, cfgs, r, out)
    finally:
        srv.close()
    srv = make_server(cfgs, capacity=2, prec=32)
    try:
        z = np.zeros_like(reqs[0]["src"])
        out, st, res, lat = srv.wait(srv.submit(0, z, 0.1, 1e-10, EVEN))
        assert st == 0 and not out.any() and res.final_rsq == 0
        r = reqs[1]
        out, st, res, lat = srv.wait(srv.submit(r["cfg"], r["src"], r["mass"], r["tol"], r["eo"]))
        assert st == 0
        check_resid(oracle, cfgs, r, out)
    finally:
        srv.close()
