"""Writes tests/golden/force_4x4x4x8.npz: seeded inputs and the oracle's
asqtad force (MILC coefficients) and HISQ force (MILC HISQ coefficients, two
epsilon-Naik groups) on them: per-direction checksums and the full matrices
at a few sites.  The reference ships no force vectors; these are generated
from the oracle after tests/test_oracle_force.py and
tests/test_oracle_hisq_force.py pinned it by finite differences, and they
freeze the results so that a later change to the generator, the oracle or
the GPU library that alters a force is caught.

Run from the repository root:  python tests/golden/make_golden_force.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import qopqdp_amd as q  # noqa: E402  (host-side generator only)
from _oracle import EVENODD, Oracle  # noqa: E402

MILC = dict(one_link=5.0 / 8, three_staple=-1.0 / 16, five_staple=1.0 / 64,
            seven_staple=-1.0 / 384, lepage=-1.0 / 16, naik=-1.0 / 24)
FAT7 = dict(one_link=1 / 8, three_staple=-1 / 16, five_staple=1 / 64, seven_staple=-1 / 384,
            lepage=0.0, naik=0.0)
LEV2 = dict(one_link=1.0, three_staple=-1 / 16, five_staple=1 / 64, seven_staple=-1 / 384,
            lepage=-1 / 8, naik=-1 / 24)

L = [4, 4, 4, 8]
GAUGE_SEED, EPS = 31, 0.3
SRC_SEEDS = [41, 42, 43]
EPS_TERMS = [0.5, -0.25, 0.125]
EPS_NAIK = [0.0, 0.05]
GROUPS = [2, 1]
SAMPLE_SITES = [0, 101, 257, 511]


def sha(arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def inputs():
    U = q.gen_gauge(L, GAUGE_SEED, EPS)
    xs = [q.gen_source(L, s, EVENODD) for s in SRC_SEEDS]
    return U, xs


def summary(F):
    s = np.array([[f[..., 0].sum(), f[..., 1].sum()] for f in F])
    n = np.array([np.sqrt((f * f).sum()) for f in F])
    samp = np.array([[f[i] for i in SAMPLE_SITES] for f in F])
    return s, n, samp


def generate():
    o = Oracle()
    U, xs = inputs()
    out = dict(L=np.array(L), gauge_seed=np.array(GAUGE_SEED), eps=np.array(EPS),
               src_seeds=np.array(SRC_SEEDS), eps_terms=np.array(EPS_TERMS),
               eps_naik=np.array(EPS_NAIK), groups=np.array(GROUPS),
               sample_sites=np.array(SAMPLE_SITES))
    out["inputs_sha256"] = np.array(sha(U + xs))
    o.rephase_milc(L, U)
    Fa = o.asqtad_force(L, U, MILC, EPS_TERMS, xs)
    out["asqtad_sum"], out["asqtad_norm"], out["asqtad_sample"] = summary(Fa)
    Fh, nf = o.hisq_force(L, U, FAT7, LEV2, 0.0, -1 / 24, EPS_NAIK, EPS_TERMS, xs, GROUPS,
                          force_filter=5e-5)
    out["hisq_sum"], out["hisq_norm"], out["hisq_sample"] = summary(Fh)
    out["hisq_filtered"] = np.array(nf)
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "force_4x4x4x8.npz"), **generate())
    print("wrote tests/golden/force_4x4x4x8.npz")
