"""Pins of the oracle's HISQ fermion force (oracle/force_oracle.c,
or_hisq_force_multi, lib/hisq_force_fnmat_p.c:7-231) and of the derivative of
the U(3) projection it uses (oracle/hisq_oracle.c, or_u3_deriv_contract,
lib/hisq_reunit1_p.c:366-433).

* Projection: Y(n,m) = dPhi/dV_mn with Phi = tr(W F^+) + tr(W^+ F),
  W = V (V^+V)^-1/2, so dPhi = 2 Re tr(dV Y) along any complex direction dV;
  checked by central differences of Phi with W from the oracle's Jacobi polar
  factor, on generic and on near-unitary V (nearly equal eigenvalues), and
  with the force filter forced on against W = V (V^+V + f)^-1/2 from numpy.
* Force: the derivative of S(U) = sum_k sum_{i in group k} eps_i
  |D_oe(fn_k(U)) X_i|^2 through fat7, the U(3) projection and the level-2
  smearing (fn_0 = asqtad(W), fn_k = fn_0 + eps_naik[k] difference(W)):
  dS/dt = Im tr(T F_mu(z)) / 2 for U_mu(z) -> exp(i t T) U_mu(z), on random
  sites, directions and generators, with one and with two epsilon-Naik
  groups.
"""
import numpy as np
import pytest

from conftest import from_raw, gauge_numpy, to_raw

EVEN, ODD = 0, 1
L = [4, 4, 4, 4]
FAT7 = dict(one_link=1 / 8, three_staple=-1 / 16, five_staple=1 / 64, seven_staple=-1 / 384,
            lepage=0.0, naik=0.0)
LEV2 = dict(one_link=1.0, three_staple=-1 / 16, five_staple=1 / 64, seven_staple=-1 / 384,
            lepage=-1 / 8, naik=-1 / 24)
DIFF_ONE, DIFF_NAIK = 0.0, -1 / 24


def cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def phi(oracle, V, F, filt=0.0):
    if filt > 0:
        w, e = np.linalg.eigh(np.conj(V.T) @ V)
        W = V @ ((e / np.sqrt(w + filt)) @ np.conj(e.T))
    else:
        W = oracle.u3_polar_jacobi(V)
    return 2 * np.real(np.trace(W @ np.conj(F.T)))


@pytest.mark.parametrize("kind", ["generic", "near_unitary", "filtered"])
def test_projection_derivative(oracle, kind):
    rng = np.random.default_rng({"generic": 1, "near_unitary": 2, "filtered": 3}[kind])
    for trial in range(5):
        if kind == "near_unitary":
            q, _ = np.linalg.qr(cplx(rng, 3, 3))
            V = 1.3 * q + 1e-4 * cplx(rng, 3, 3)
        else:
            V = cplx(rng, 3, 3) + (2 * np.eye(3) if kind == "generic" else 0)
        F = cplx(rng, 3, 3)
        filt = 0.0
        if kind == "filtered":
            lmin = np.linalg.eigvalsh(np.conj(V.T) @ V).min()
            filt = 2 * lmin + 0.1  # acts
        Y, k = oracle.u3_deriv_contract(V, F, filt)
        assert k == (1 if kind == "filtered" else 0)
        for d in range(3):
            dV = cplx(rng, 3, 3)
            h = 1e-5
            fd = (phi(oracle, V + h * dV, F, filt) - phi(oracle, V - h * dV, F, filt)) / (2 * h)
            pred = 2 * np.real(np.trace(dV @ Y))
            assert abs(fd - pred) <= 1e-7 * max(1.0, abs(pred)), (kind, trial, fd, pred)


def hisq_links(oracle, U, eps_naik):
    W, _, _ = oracle.hisq_W(L, U, FAT7)  # the Jacobi fallback, where taken, is the same polar factor
    fat, lng = oracle.make_imp_links(L, W, LEV2)
    sets = [(fat, lng)]
    if len(eps_naik) > 1:
        f3, l3 = oracle.make_imp_links(L, W, dict(one_link=DIFF_ONE, three_staple=0.0, five_staple=0.0,
                                                  seven_staple=0.0, lepage=0.0, naik=DIFF_NAIK),
                                       with_long=True)
        for e in eps_naik[1:]:
            sets.append(([a + e * b for a, b in zip(fat, f3)], [a + e * b for a, b in zip(lng, l3)]))
    return sets


def d_oe(oracle, fl, X):
    h = oracle.links(L, *fl)
    try:
        return oracle.dslash(h, 0.0, X, ODD, EVEN)
    finally:
        oracle.free_links(h)


@pytest.mark.parametrize("groups", [(2,), (2, 1)])
def test_hisq_force_is_the_derivative(oracle, groups):
    eps_naik = [0.0, 0.1][:len(groups)]
    # phase-absorbed links, as the HISQ links are built from: the staggered
    # phases make the fat7 free-field weight 1 (without them it is 0 and the
    # projection is singular)
    U = gauge_numpy(L, 9300 + len(groups), 0.4)
    oracle.rephase_milc(L, U)
    V = int(np.prod(L))
    rng = np.random.default_rng(9400 + len(groups))
    n = sum(groups)
    Xs = []
    for _ in range(n):
        X = rng.standard_normal((V, 3, 2))
        X[V // 2:] = 0
        Xs.append(X)
    eps = [0.8, -0.3, 0.5][:n]
    gid = [k for k, c in enumerate(groups) for _ in range(c)]

    def action(Uc):
        sets = hisq_links(oracle, Uc, eps_naik)
        return sum(e * float(np.sum(d_oe(oracle, sets[k], X)[V // 2:] ** 2))
                   for e, X, k in zip(eps, Xs, gid))

    sets = hisq_links(oracle, U, eps_naik)
    xs = []
    for X, k in zip(Xs, gid):
        x = X.copy()
        x[V // 2:] = d_oe(oracle, sets[k], X)[V // 2:]
        xs.append(x)
    F, nf = oracle.hisq_force(L, U, FAT7, LEV2, DIFF_ONE, DIFF_NAIK, eps_naik, eps, xs, list(groups),
                              force_filter=5e-5)
    assert nf == 0
    for mu in range(4):
        f = from_raw(F[mu])
        assert np.abs(f + np.conj(np.transpose(f, (0, 2, 1)))).max() < 1e-12
        assert np.abs(np.trace(f, axis1=1, axis2=2)).max() < 1e-12
    h = 1e-4
    for _ in range(5):
        z = int(rng.integers(V))
        mu = int(rng.integers(4))
        g = cplx(rng, 3, 3)
        T = 0.5 * (g + np.conj(g.T))
        T -= np.trace(T) / 3 * np.eye(3)
        w, v = np.linalg.eigh(T)
        u0 = from_raw(U[mu][z])
        s = []
        for t in (h, -h):
            Up = [u.copy() for u in U]
            Up[mu][z] = to_raw((((v * np.exp(1j * t * w)) @ np.conj(v.T)) @ u0)[None])[0]
            s.append(action(Up))
        fd = (s[0] - s[1]) / (2 * h)
        pred = 0.5 * np.imag(np.trace(T @ from_raw(F[mu][z])))
        assert abs(fd - pred) <= 1e-6 * max(abs(pred), 1e-3) + 1e-9, (groups, z, mu, fd, pred)
