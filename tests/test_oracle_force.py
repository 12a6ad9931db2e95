"""The oracle's asqtad fermion force (lib/asqtad_force_fnmat_p.c:474-546,
oracle/force_oracle.c) against the derivative it must be.

With x_i = X_i on even sites and Y_i = D_oe X_i on odd sites, the force of
sum_i eps_i x_i x_i^+ is the derivative of S(U) = sum_i eps_i |D_oe(U) X_i|^2,
where D_oe is built from the fat and Naik links smeared from U.  For
U_mu(z) -> exp(i t T) U_mu(z) with T Hermitian and traceless,

    dS/dt = Im tr(T F_mu(z)) / 2,

F the (traceless antihermitian) force.  Central differences check this on
random sites, directions and generators, for the MILC coefficients and for
coefficient sets that switch on one smearing term at a time (one-link,
3-staple, 5-staple, 7-staple, Lepage, Naik), so a wrong term, sign or shift
fails by itself."""
import numpy as np
import pytest

from conftest import MILC, from_raw, gauge_numpy, to_raw

EVEN, ODD = 0, 1
L = [4, 4, 4, 4]

ZERO = dict(one_link=0.0, three_staple=0.0, five_staple=0.0, seven_staple=0.0, lepage=0.0, naik=0.0)
SETS = {
    "milc": MILC,
    "one_link": dict(ZERO, one_link=0.7),
    "three_staple": dict(ZERO, three_staple=0.3),
    "five_staple": dict(ZERO, five_staple=0.2),
    "seven_staple": dict(ZERO, seven_staple=0.2),
    "lepage": dict(ZERO, lepage=0.25),
    "naik": dict(ZERO, one_link=0.1, naik=0.4),
}


def action(oracle, U, coeffs, eps, Xs):
    fat, lng = oracle.make_imp_links(L, U, coeffs)
    h = oracle.links(L, fat, lng)
    try:
        s = 0.0
        for e, X in zip(eps, Xs):
            Y = oracle.dslash(h, 0.0, X, ODD, EVEN)
            s += e * float(np.sum(Y[len(Y) // 2:] ** 2))
        return s
    finally:
        oracle.free_links(h)


def full_vectors(oracle, U, coeffs, Xs):
    fat, lng = oracle.make_imp_links(L, U, coeffs)
    h = oracle.links(L, fat, lng)
    try:
        out = []
        for X in Xs:
            Y = oracle.dslash(h, 0.0, X, ODD, EVEN)
            x = X.copy()
            x[len(x) // 2:] = Y[len(Y) // 2:]
            out.append(x)
        return out
    finally:
        oracle.free_links(h)


def expm_i(T, t):
    w, v = np.linalg.eigh(T)
    return (v * np.exp(1j * t * w)) @ np.conj(v.T)


@pytest.mark.parametrize("name", list(SETS))
def test_force_is_the_derivative_of_the_action(oracle, name):
    coeffs = SETS[name]
    U = gauge_numpy(L, 7100, 0.5)
    V = int(np.prod(L))
    rng = np.random.default_rng(7200)
    Xs = []
    for _ in range(2):
        X = rng.standard_normal((V, 3, 2))
        X[V // 2:] = 0
        Xs.append(X)
    eps = [0.7, -0.4]
    xs = full_vectors(oracle, U, coeffs, Xs)
    F = oracle.asqtad_force(L, U, coeffs, eps, xs)
    # the force is traceless antihermitian
    for mu in range(4):
        f = from_raw(F[mu])
        assert np.abs(f + np.conj(np.transpose(f, (0, 2, 1)))).max() < 1e-12
        assert np.abs(np.trace(f, axis1=1, axis2=2)).max() < 1e-12
    h = 1e-4
    for k in range(6):
        z = int(rng.integers(V))
        mu = int(rng.integers(4))
        g = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        T = 0.5 * (g + np.conj(g.T))
        T -= np.trace(T) / 3 * np.eye(3)
        u0 = from_raw(U[mu][z])
        s = []
        for t in (h, -h):
            Up = [u.copy() for u in U]
            Up[mu][z] = to_raw((expm_i(T, t) @ u0)[None])[0]
            s.append(action(oracle, Up, coeffs, eps, Xs))
        fd = (s[0] - s[1]) / (2 * h)
        pred = 0.5 * np.imag(np.trace(T @ from_raw(F[mu][z])))
        scale = max(abs(pred), 1e-3)
        assert abs(fd - pred) <= 1e-6 * scale + 1e-9, (name, z, mu, fd, pred)


def test_force_accumulates_and_is_linear_in_eps(oracle):
    U = gauge_numpy(L, 7300, 0.4)
    V = int(np.prod(L))
    rng = np.random.default_rng(7400)
    x = rng.standard_normal((V, 3, 2))
    F1 = oracle.asqtad_force(L, U, MILC, [1.0], [x])
    F2 = oracle.asqtad_force(L, U, MILC, [2.5], [x], force=[f.copy() for f in F1])
    for mu in range(4):
        np.testing.assert_allclose(F2[mu], 3.5 * F1[mu], rtol=1e-12, atol=1e-13)
