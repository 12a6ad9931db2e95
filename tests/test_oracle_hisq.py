"""Pins of the oracle's HISQ link preparation (oracle/hisq_oracle.c):
U(3) projection W = V (V^dagger V)^-1/2 (lib/hisq_reunit_utilities_p.c:97-297)
against numpy's SVD, an independent eigen path and closed forms, and the
fat7 -> U(3) stage (lib/hisq_links_p.c:62-152) in the free field and under
gauge transformations."""
import numpy as np

from conftest import from_raw, gauge_numpy, rand_su3, to_raw
from test_oracle import coords_all, free_field, gauge_transform, plane_wave, shift

EVEN = 0
# MILC HISQ coefficients: fat7 level and the asqtad level 2 (epsilon_Naik = 0)
HISQ_FAT7 = dict(one_link=1 / 8, three_staple=-1 / 16, five_staple=1 / 64,
                 seven_staple=-1 / 384, lepage=0.0, naik=0.0)
HISQ_L2 = dict(one_link=1.0, three_staple=-1 / 16, five_staple=1 / 64,
               seven_staple=-1 / 384, lepage=-1 / 8, naik=-1 / 24)


def polar_np(V):
    u, s, vh = np.linalg.svd(V)
    return u @ vh


def rand_mats(n, seed, shift_=3.0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2) \
        + shift_ * np.eye(3)


def test_u3_analytic_is_the_polar_factor(oracle):
    for V in rand_mats(200, 1):
        W, svd, fl = oracle.u3_un_analytic(V)
        assert svd == 0
        # the cubic loses a few digits on generic (not near-unitary) matrices
        assert np.abs(W.conj().T @ W - np.eye(3)).max() < 1e-11
        H = W.conj().T @ V  # = (V^dagger V)^1/2: hermitian positive definite
        assert np.abs(H - H.conj().T).max() < 1e-12
        assert np.linalg.eigvalsh(0.5 * (H + H.conj().T)).min() > 0
        np.testing.assert_allclose(W, polar_np(V), atol=1e-11)
        np.testing.assert_allclose(oracle.u3_polar_jacobi(V), polar_np(V), atol=1e-12)
        assert fl == 795  # flops of the analytic path, as the reference counts them


def test_u3_scaled_unitary_returns_the_unitary(oracle):
    """V = c U: equal eigenvalues of V^dagger V (the |S| < eps branch), W = U."""
    for i, U in enumerate(rand_su3(20, np.random.default_rng(2))):
        U = U * np.exp(0.3j * i)  # a U(3) phase as well
        W, svd, _ = oracle.u3_un_analytic(1.7 * U)
        assert svd == 0
        np.testing.assert_allclose(W, U, atol=1e-14)


def test_u3_svd_only_and_ill_conditioned_fallback(oracle):
    o = oracle.reunit_opts(svd_only=1)
    for V in rand_mats(50, 3):
        Wa, _, _ = oracle.u3_un_analytic(V)
        Ws, svd, _ = oracle.u3_un_analytic(V, o)
        assert svd == 1
        np.testing.assert_allclose(Ws, Wa, atol=1e-11)
    # |det V|^2 below svd_abs_error: the fallback is taken and still returns the polar factor
    rng = np.random.default_rng(4)
    for _ in range(20):
        u = rand_su3(1, rng)[0]
        v = rand_su3(1, rng)[0]
        V = u @ np.diag([1.0, 0.5, 1e-5]) @ v
        W, svd, _ = oracle.u3_un_analytic(V)
        assert svd == 1
        np.testing.assert_allclose(W, polar_np(V), atol=1e-9)
    # with the fallback disabled the analytic path is used whatever V is
    W, svd, _ = oracle.u3_un_analytic(V, oracle.reunit_opts(allow_svd=0))
    assert svd == 0


def test_hisq_W_free_field(oracle):
    """U = 1 with MILC phases and antiperiodic t: the fat7 weight is 1, so
    W = U; level 2 then has weights 9/8 and -1/24 and A has the asqtad
    free spectrum."""
    L = [4, 4, 4, 8]
    U = free_field(L)
    oracle.rephase_milc(L, U)
    W, nsvd, _ = oracle.hisq_W(L, U, HISQ_FAT7)
    assert nsvd == 0
    for mu in range(4):
        np.testing.assert_allclose(W[mu], U[mu], atol=1e-14)
    fat, lng = oracle.make_imp_links(L, W, HISQ_L2)
    h = oracle.links(L, fat, lng)
    V = int(np.prod(L))
    m = 0.05
    for n in ([0, 0, 0, 0], [1, 2, 0, 3], [3, 3, 1, 6]):
        v, p = plane_wave(L, n)
        s = 9 / 8 * np.sin(p) - 1 / 24 * np.sin(3 * p)
        Av, _ = oracle.ddagd_norm2(h, m, v, EVEN)
        np.testing.assert_allclose(Av[: V // 2], (m * m + np.sum(s * s)) * v[: V // 2], atol=1e-12)
    oracle.free_links(h)


def test_hisq_W_is_polar_of_fat7_and_covariant(oracle):
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    U = gauge_numpy(L, 31)
    oracle.rephase_milc(L, U)
    W, nsvd, fl = oracle.hisq_W(L, U, HISQ_FAT7)
    assert nsvd == 0 and fl == 795 * 4 * V
    fat7, _ = oracle.make_imp_links(L, U, HISQ_FAT7)
    for mu in range(4):
        w = from_raw(W[mu])
        np.testing.assert_allclose(w, polar_np(from_raw(fat7[mu])), atol=1e-13)
        # fat7 links of a smooth field are near-unitary: W is U(3) to rounding
        ww = np.einsum("nki,nkj->nij", w.conj(), w)
        assert np.abs(ww - np.eye(3)).max() < 1e-14
    g = rand_su3(V, np.random.default_rng(32))
    Wg, _, _ = oracle.hisq_W(L, gauge_transform(L, U, g), HISQ_FAT7)
    x = coords_all(L)
    for mu in range(4):
        gp = g[shift(L, x, mu, 1)]
        exp = g @ from_raw(W[mu]) @ np.conj(np.transpose(gp, (0, 2, 1)))
        np.testing.assert_allclose(from_raw(Wg[mu]), exp, atol=1e-12)
