"""Pins of the oracle's eigCG (oracle/qop_oracle.c invert_eigcg, restating
lib/invert_eigcg_p.c:1-1068 with Jacobi in place of LAPACK's dsyev/zheev),
driven through the EVENODD invert with cgtype = 1
(lib/asqtad_invert_p.c:683-687, 767-769).

* the Ritz values of the deflation space converge to the lowest eigenvalues
  of A = m^2 - D_eo D_oe, computed densely on 4x4x4x8 (1e-6 relative);
* every solve reaches its residual and equals the plain CG solution;
* deflation cuts the iterations of later solves;
* a space that fills up stops growing (no vectors beyond numax), and a solve
  on the other parity starts a new space instead of producing NaNs.
"""
import numpy as np
import pytest

from conftest import MILC, gauge_numpy

L = [4, 4, 4, 8]
V = 512
MASS = 0.003


@pytest.fixture(scope="module")
def setup(oracle):
    U = gauge_numpy(L, 11, 0.3)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    h = oracle.links(L, fat, lng)
    yield oracle, h
    oracle.free_links(h)


def dense_A(oracle, h):
    n = V // 2 * 3
    A = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros((V, 3, 2))
        e[k // 3, k % 3, 0] = 1
        out, _ = oracle.ddagd_norm2(h, MASS, e, 0)
        A[:, k] = (out[:V // 2, :, 0] + 1j * out[:V // 2, :, 1]).reshape(-1)
    return A


def sources(seed, n, even_heavy=True):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        b = rng.standard_normal((V, 3, 2))
        if even_heavy:
            b[V // 2:] *= 0.5  # the even part dominates: the CG runs on even sites
        out.append(b)
    return out


def test_eigcg_ritz_values_solutions_and_iterations(setup):
    oracle, h = setup
    w = np.linalg.eigvalsh(dense_A(oracle, h))
    eg = oracle.eigcg_state()
    try:
        iters = []
        for b in sources(1, 8):
            x, ra = oracle.invert_eigcg(h, eg, MASS, b, 1e-20, numax=100, m=24, nev=6)
            xc, rc, _ = oracle.invert(h, MASS, b, 2, 1e-20)
            assert ra.final_rsq <= 1e-20
            assert np.abs(x - xc).max() <= 1e-8 * np.abs(xc).max()
            iters.append(ra.final_iter)
            assert rc.final_iter >= 115  # plain CG: 119-122 iterations
        ritz = oracle.eigcg_ritz(eg)
        assert len(ritz) >= 8
        np.testing.assert_allclose(ritz[:4], w[:4], rtol=1e-6)
        assert np.all(np.diff(ritz) >= -1e-12)  # ascending
        assert iters[-1] <= 0.9 * iters[0], iters
    finally:
        oracle.eigcg_free(eg)


def test_eigcg_space_bounds_and_parity_change(setup):
    oracle, h = setup
    eg = oracle.eigcg_state()
    try:
        nus = []
        for b in sources(2, 6):
            x, ra = oracle.invert_eigcg(h, eg, MASS, b, 1e-16, numax=40, m=24, nev=6)
            assert ra.final_rsq <= 1e-16 and np.isfinite(x).all()
            nus.append(len(oracle.eigcg_ritz(eg)))
        # grows by nev while nu + m fits in numax, then stops
        assert nus == [6, 12, 18, 18, 18, 18], nus
        # odd-dominated source: the CG runs on odd sites; a new space starts
        b = sources(3, 1, even_heavy=False)[0]
        b[:V // 2] *= 0.1
        x, ra = oracle.invert_eigcg(h, eg, MASS, b, 1e-16, numax=40, m=24, nev=6)
        assert ra.final_rsq <= 1e-16 and np.isfinite(x).all()
        assert np.isfinite(oracle.eigcg_ritz(eg)).all()
    finally:
        oracle.eigcg_free(eg)
