"""Writes tests/golden/asqtad_4x4x4x8.npz: seeded inputs, their checksums
and the oracle's outputs on them (fattened-link checksums, Dslash, EVEN
solve, multi-mass solve).  The reference ships no golden vectors (its only
known-answer check is check_op in examples/asqtad_test.c:37-81), so these
fixtures are generated from the oracle once, after the oracle itself was
checked against the closed forms and invariants in tests/test_oracle.py;
they then freeze the inputs and outputs so that any later change to the
generator, the oracle or the GPU library that alters a result is caught.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import qopqdp_amd as q  # noqa: E402  (host-side generator only)
from _oracle import EVEN, EVENODD, Oracle  # noqa: E402

MILC = dict(one_link=5.0 / 8, three_staple=-1.0 / 16, five_staple=1.0 / 64,
            seven_staple=-1.0 / 384, lepage=-1.0 / 16, naik=-1.0 / 24)

L = [4, 4, 4, 8]
GAUGE_SEED, EPS = 11, 0.3
SRC_SEEDS = [21, 22]
DSLASH_MASS = 0.05
SOLVE_MASS = 0.1
MULTI_MASSES = [0.05, 0.1, 0.2]
RSQMIN = 1e-24
SAMPLE_SITES = [0, 77, 300]


def sha(arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def link_checksums(links):
    s = np.array([[l[..., 0].sum(), l[..., 1].sum()] for l in links])
    n = np.array([np.sqrt((l * l).sum()) for l in links])
    return s, n


def generate():
    o = Oracle()
    U = q.gen_gauge(L, GAUGE_SEED, EPS)
    out = dict(L=np.array(L), gauge_seed=np.array(GAUGE_SEED), eps=np.array(EPS),
               src_seeds=np.array(SRC_SEEDS), dslash_mass=np.array(DSLASH_MASS),
               solve_mass=np.array(SOLVE_MASS), multi_masses=np.array(MULTI_MASSES),
               rsqmin=np.array(RSQMIN), sample_sites=np.array(SAMPLE_SITES))
    out["gauge_sha256"] = np.array(sha(U))
    out["plaquette"] = np.array(o.plaquette(L, U))
    o.rephase_milc(L, U)
    out["rephased_sha256"] = np.array(sha(U))
    fat, lng = o.make_imp_links(L, U, MILC)
    out["fat_sum"], out["fat_norm"] = link_checksums(fat)
    out["long_sum"], out["long_norm"] = link_checksums(lng)
    out["fat_sites"] = np.stack([f[SAMPLE_SITES] for f in fat])
    out["long_sites"] = np.stack([l[SAMPLE_SITES] for l in lng])
    h = o.links(L, fat, lng)
    srcs = [q.gen_source(L, s, EVENODD) for s in SRC_SEEDS]
    out["src_sha256"] = np.array([sha([s]) for s in srcs])
    out["dslash"] = np.stack([o.dslash(h, DSLASH_MASS, s, EVENODD, EVENODD) for s in srcs])
    b = srcs[0].copy()
    b[b.shape[0] // 2:] = 0
    x, ra, _ = o.invert(h, SOLVE_MASS, b, EVEN, RSQMIN)
    out["solve_x"] = x
    out["solve_iters"] = np.array(ra.final_iter)
    out["solve_rsq"] = np.array(ra.final_rsq)
    outs, ras, _ = o.invert_multi(h, [MULTI_MASSES], [b], EVEN, RSQMIN)
    out["multi_x"] = np.stack(outs)
    out["multi_iters"] = np.array([r.final_iter for r in ras])
    o.free_links(h)
    return out


if __name__ == "__main__":
    data = generate()
    path = os.path.join(HERE, "asqtad_4x4x4x8.npz")
    np.savez_compressed(path, **data)
    print("wrote", path, os.path.getsize(path), "bytes")
