"""Pins for the CPU oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a wrong sign,
a transposed link, an index off by one, a missing factor 1/2) fails it, and
none of them retypes the oracle's own formula:
  * geometry: raw index <-> coordinates bijection, even sites first;
  * rephase_G: closed form per site;
  * Dslash: agreement with an independent check_op restatement
    (examples/asqtad_test.c:37-81) and with a dense operator assembled here
    from the link fields with numpy; anti-hermiticity D_eo = -D_oe^+;
    gauge covariance;
  * fattening: one-link-only and Naik-only closed forms, free-field weights
    f = 9/8 and l = -1/24 with MILC phases and antiperiodic time, gauge
    covariance;
  * A = m^2 - D_eo D_oe: free-field plane-wave spectrum m^2 + sum_mu s_mu^2;
  * solvers: exact free-field solution, dense-solve agreement, the
    check_resids identity, multi-mass vs single-mass.
"""
import numpy as np
import pytest

from conftest import MILC, from_raw, gauge_numpy, rand_su3, to_raw

EVEN, ODD, EVENODD = 0, 1, 2


def coords_all(L):
    V = int(np.prod(L))
    lex = np.arange(V)
    x = np.stack([lex // (L[1] * L[2] * L[3]), (lex // (L[2] * L[3])) % L[1],
                  (lex // L[3]) % L[2], lex % L[3]], axis=1)
    par = x.sum(axis=1) & 1
    idx = np.where(par == 0, lex // 2, (lex + V) // 2)
    out = np.zeros_like(x)
    out[idx] = x
    return out


def index_of(L, x):
    V = int(np.prod(L))
    lex = ((x[:, 0] * L[1] + x[:, 1]) * L[2] + x[:, 2]) * L[3] + x[:, 3]
    return np.where((x.sum(axis=1) & 1) == 0, lex // 2, (lex + V) // 2)


def shift(L, x, mu, d):
    y = x.copy()
    y[:, mu] = (y[:, mu] + d) % L[mu]
    return index_of(L, y)


def random_links(L, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(L))
    f = [np.ascontiguousarray(rng.standard_normal((V, 3, 3, 2))) for _ in range(4)]
    l = [np.ascontiguousarray(rng.standard_normal((V, 3, 3, 2))) for _ in range(4)]
    return f, l


def rand_vec(L, seed, eo=EVENODD):
    rng = np.random.default_rng(seed)
    V = int(np.prod(L))
    v = rng.standard_normal((V, 3, 2))
    if eo == EVEN:
        v[V // 2:] = 0
    elif eo == ODD:
        v[: V // 2] = 0
    return v


def test_geometry_roundtrip(oracle):
    L = [4, 6, 4, 8]
    V = int(np.prod(L))
    x = coords_all(L)
    assert len({tuple(r) for r in x}) == V
    for i in range(0, V, 37):
        assert oracle.raw_coords(L, i) == list(x[i])
        assert oracle.raw_index(L, list(x[i])) == i
    par = x.sum(axis=1) & 1
    assert np.all(par[: V // 2] == 0) and np.all(par[V // 2:] == 1)


def test_rephase_closed_form(oracle):
    L = [4, 4, 6, 8]
    V = int(np.prod(L))
    eye = np.zeros((V, 3, 3, 2))
    eye[:, [0, 1, 2], [0, 1, 2], 0] = 1
    U = [eye.copy() for _ in range(4)]
    r0 = [0, 1, 0, 2]
    oracle.rephase(L, U, r0, [1, 1, 1j, -1], [0, 1, 3, 7])
    x = coords_all(L)
    for mu, mask, bc in [(0, 0, 1), (1, 1, 1), (2, 3, 1j), (3, 7, -1)]:
        s = sum(((mask >> j) & 1) * x[:, j] for j in range(4)) & 1
        eta = np.where(s == 1, -1.0, 1.0)
        last = (x[:, mu] + L[mu] - r0[mu]) % L[mu] == L[mu] - 1
        ph = eta * np.where(last, bc, 1)
        got = from_raw(U[mu])
        np.testing.assert_allclose(got, ph[:, None, None] * np.eye(3), atol=0)


def test_dslash_matches_check_op(oracle):
    L = [4, 4, 4, 6]
    fat, lng = random_links(L, 1)
    h = oracle.links(L, fat, lng)
    v = rand_vec(L, 2)
    for mass in (0.0, 0.37):
        got = oracle.dslash(h, mass, v, EVENODD, EVENODD)
        ref = oracle.check_op(L, fat, lng, mass, v)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-13
    # without long links (nlinks = 8)
    h8 = oracle.links(L, fat, None)
    got = oracle.dslash(h8, 0.2, v, EVENODD, EVENODD)
    ref = oracle.check_op(L, fat, None, 0.2, v)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13
    oracle.free_links(h)
    oracle.free_links(h8)


def dense_hop(L, fat, lng):
    """Dense hop matrix assembled from the user links:
    D x = 1/2 sum_mu [F(x) v(x+mu) - F(x-mu)^+ v(x-mu) + L(x) v(x+3mu) - L(x-3mu)^+ v(x-3mu)]."""
    V = int(np.prod(L))
    x = coords_all(L)
    D = np.zeros((V * 3, V * 3), dtype=complex)
    ar = np.arange(V)
    for mu in range(4):
        F = from_raw(fat[mu])
        Lk = from_raw(lng[mu])
        for d, M in ((1, F), (3, Lk)):
            fwd = shift(L, x, mu, d)
            bwd = shift(L, x, mu, -d)
            for a in range(3):
                for b in range(3):
                    D[3 * ar + a, 3 * fwd + b] += 0.5 * M[ar, a, b]
                    D[3 * ar + a, 3 * bwd + b] -= 0.5 * np.conj(M[bwd, b, a])
    return D


def test_dslash_dense_and_antihermitian(oracle):
    L = [4, 4, 4, 4]
    V = int(np.prod(L))
    fat, lng = random_links(L, 3)
    h = oracle.links(L, fat, lng)
    D = dense_hop(L, fat, lng)
    # apply the oracle's D_oe (out odd, in even) to random vectors and compare
    v = rand_vec(L, 4, EVEN)
    got = oracle.dslash(h, 0.0, v, ODD, EVEN)
    vc = (v[..., 0] + 1j * v[..., 1]).reshape(-1)
    ref = (D @ vc).reshape(V, 3)
    gc = got[..., 0] + 1j * got[..., 1]
    np.testing.assert_allclose(gc[V // 2:], ref[V // 2:], rtol=0, atol=1e-12)
    # hop is anti-hermitian
    assert np.abs(D + D.conj().T).max() < 1e-13
    # even output ignores the odd part of the output vector (only subset written)
    out = np.full_like(v, 7.0)
    oracle.dslash(h, 0.0, v, ODD, EVEN, out=out)
    assert np.all(out[: V // 2] == 7.0)
    oracle.free_links(h)


def test_dslash_eo_dispatch(oracle):
    """D(eo_out, eo_in) blocks of the full operator (asqtad_dslash_p.c:1048-1092)."""
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    fat, lng = random_links(L, 5)
    h = oracle.links(L, fat, lng)
    m = 0.3
    v = rand_vec(L, 6)
    full = oracle.dslash(h, m, v, EVENODD, EVENODD)
    ve = v.copy(); ve[V // 2:] = 0
    vo = v.copy(); vo[: V // 2] = 0
    e, o = slice(0, V // 2), slice(V // 2, V)
    De_e = oracle.dslash(h, m, v, EVEN, EVEN)
    De_o = oracle.dslash(h, m, v, EVEN, ODD)
    Do_e = oracle.dslash(h, m, v, ODD, EVEN)
    Do_o = oracle.dslash(h, m, v, ODD, ODD)
    np.testing.assert_allclose(De_e[e], m * v[e], atol=0)
    np.testing.assert_allclose(Do_o[o], m * v[o], atol=0)
    np.testing.assert_allclose(De_e[e] + De_o[e], full[e], atol=1e-12)
    np.testing.assert_allclose(Do_e[o] + Do_o[o], full[o], atol=1e-12)
    De_all = oracle.dslash(h, m, v, EVEN, EVENODD)
    np.testing.assert_allclose(De_all[e], full[e], atol=1e-12)
    oracle.free_links(h)


def gauge_transform(L, U, g):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^+ ; g (V,3,3) complex."""
    x = coords_all(L)
    out = []
    for mu in range(4):
        u = from_raw(U[mu])
        gp = g[shift(L, x, mu, 1)]
        out.append(to_raw(g @ u @ np.conj(np.transpose(gp, (0, 2, 1)))))
    return out


def test_fattening_gauge_covariance(oracle):
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    U = gauge_numpy(L, 11)
    oracle.rephase_milc(L, U)
    g = rand_su3(V, np.random.default_rng(12))
    Ug = gauge_transform(L, U, g)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    fatg, lngg = oracle.make_imp_links(L, Ug, MILC)
    x = coords_all(L)
    for mu in range(4):
        for M, Mg, d in ((fat[mu], fatg[mu], 1), (lng[mu], lngg[mu], 3)):
            gp = g[shift(L, x, mu, d)]
            exp = g @ from_raw(M) @ np.conj(np.transpose(gp, (0, 2, 1)))
            np.testing.assert_allclose(from_raw(Mg), exp, atol=1e-12)
    # and the Dslash is covariant: D[U^g](g psi) = g D[U] psi
    h, hg = oracle.links(L, fat, lng), oracle.links(L, fatg, lngg)
    v = rand_vec(L, 13)
    vc = from_raw(v)
    gv = to_raw(np.einsum("nab,nb->na", g, vc))
    lhs = oracle.dslash(hg, 0.1, gv, EVENODD, EVENODD)
    rhs = to_raw(np.einsum("nab,nb->na", g, from_raw(oracle.dslash(h, 0.1, v, EVENODD, EVENODD))))
    assert np.abs(lhs - rhs).max() < 1e-12
    oracle.free_links(h)
    oracle.free_links(hg)


def test_fattening_closed_forms(oracle):
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    U = gauge_numpy(L, 21)
    x = coords_all(L)
    one = dict(one_link=0.7, three_staple=0, five_staple=0, seven_staple=0, lepage=0, naik=0)
    fat, lng = oracle.make_imp_links(L, U, one)
    assert lng is None
    for mu in range(4):
        np.testing.assert_allclose(fat[mu], 0.7 * U[mu], atol=1e-15)
    naik = dict(one_link=0, three_staple=0, five_staple=0, seven_staple=0, lepage=0, naik=-0.25)
    fat, lng = oracle.make_imp_links(L, U, naik)
    for mu in range(4):
        u = from_raw(U[mu])
        exp = -0.25 * u @ u[shift(L, x, mu, 1)] @ u[shift(L, x, mu, 2)]
        np.testing.assert_allclose(from_raw(lng[mu]), exp, atol=1e-13)
        assert np.abs(fat[mu]).max() == 0


def free_field(L):
    V = int(np.prod(L))
    eye = np.zeros((V, 3, 3, 2))
    eye[:, [0, 1, 2], [0, 1, 2], 0] = 1
    return [eye.copy() for _ in range(4)]


def test_fattening_free_field_weights(oracle):
    """U = 1 with MILC phases and antiperiodic t: fat = 9/8 eta bc, long = -1/24 eta bc."""
    L = [4, 4, 4, 8]
    U = free_field(L)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    x = coords_all(L)
    for mu, mask in zip(range(4), (0, 1, 3, 7)):
        s = sum(((mask >> j) & 1) * x[:, j] for j in range(4)) & 1
        eta = np.where(s == 1, -1.0, 1.0)
        bcf = np.where((mu == 3) & (x[:, 3] == L[3] - 1), -1.0, 1.0)
        bcl = np.where((mu == 3) & (x[:, 3] >= L[3] - 3), -1.0, 1.0)
        np.testing.assert_allclose(from_raw(fat[mu]), (9 / 8 * eta * bcf)[:, None, None] * np.eye(3),
                                   atol=1e-14)
        np.testing.assert_allclose(from_raw(lng[mu]), (-1 / 24 * eta * bcl)[:, None, None] * np.eye(3),
                                   atol=1e-15)


def plane_wave(L, n, eo=EVEN, color=(1, 0, 0)):
    """exp(i p.x) with p_i = 2 pi n_i / L_i (i<3), p_t = (2 n_t + 1) pi / L_t."""
    x = coords_all(L)
    p = np.array([2 * np.pi * n[i] / L[i] for i in range(3)] + [(2 * n[3] + 1) * np.pi / L[3]])
    ph = np.exp(1j * (x @ p))
    V = int(np.prod(L))
    v = ph[:, None] * np.array(color, dtype=complex)[None, :]
    if eo == EVEN:
        v[V // 2:] = 0
    return to_raw(v), p


def test_free_field_spectrum(oracle):
    L = [4, 4, 4, 8]
    U = free_field(L)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    h = oracle.links(L, fat, lng)
    V = int(np.prod(L))
    m = 0.07
    for n in ([0, 0, 0, 0], [1, 0, 2, 3], [3, 1, 1, 5], [2, 2, 2, 7]):
        v, p = plane_wave(L, n)
        s = 9 / 8 * np.sin(p) - 1 / 24 * np.sin(3 * p)
        lam = m * m + np.sum(s * s)
        Av, pAp = oracle.ddagd_norm2(h, m, v, EVEN)
        np.testing.assert_allclose(Av[: V // 2], lam * v[: V // 2], atol=1e-12)
        np.testing.assert_allclose(pAp, lam * np.sum(v * v), rtol=1e-12)
    oracle.free_links(h)


def test_free_field_solution(oracle):
    """EVEN solve of a plane wave: out = m/(m^2 + sum s^2) b exactly (README:24-48)."""
    L = [4, 4, 4, 8]
    U = free_field(L)
    oracle.rephase_milc(L, U)
    fat, lng = oracle.make_imp_links(L, U, MILC)
    h = oracle.links(L, fat, lng)
    V = int(np.prod(L))
    m = 0.05
    v, p = plane_wave(L, [1, 2, 0, 1], color=(0.3, -0.2j, 1))
    s = 9 / 8 * np.sin(p) - 1 / 24 * np.sin(3 * p)
    lam = m * m + np.sum(s * s)
    out, ra, _ = oracle.invert(h, m, v, EVEN, 1e-24)
    np.testing.assert_allclose(out[: V // 2], m / lam * v[: V // 2], rtol=0, atol=1e-10)
    assert ra.final_rsq < 1e-20
    oracle.free_links(h)


def smeared_links(oracle, L, seed):
    U = gauge_numpy(L, seed)
    oracle.rephase_milc(L, U)
    return oracle.make_imp_links(L, U, MILC)


def test_even_solve_dense_and_check_resids(oracle):
    L = [4, 4, 4, 4]
    V = int(np.prod(L))
    fat, lng = smeared_links(oracle, L, 31)
    h = oracle.links(L, fat, lng)
    m = 0.2
    b = rand_vec(L, 32, EVEN)
    out, ra, flop = oracle.invert(h, m, b, EVEN, 1e-20)
    # dense reference: A = m^2 - D_eo D_oe on even sites
    D = dense_hop(L, fat, lng)
    Vh = V // 2
    ev = np.concatenate([3 * np.arange(Vh) + c for c in range(3)])
    od = np.concatenate([3 * np.arange(Vh, V) + c for c in range(3)])
    ev.sort(); od.sort()
    A = m * m * np.eye(len(ev)) - D[np.ix_(ev, od)] @ D[np.ix_(od, ev)]
    bc = from_raw(b).reshape(-1)[ev]
    x = np.linalg.solve(A, bc) * m
    np.testing.assert_allclose(from_raw(out).reshape(-1)[ev], x, atol=1e-8)
    assert oracle.true_resid_even(h, m, out, b) < 1e-9
    assert ra.final_rsq < 1e-20 * 1.01
    assert flop == pytest.approx(1182 * ra.final_iter * V)
    oracle.free_links(h)


def test_evenodd_solve_residual(oracle):
    L = [4, 4, 4, 6]
    fat, lng = smeared_links(oracle, L, 41)
    h = oracle.links(L, fat, lng)
    b = rand_vec(L, 42)
    for m in (0.5, 0.1):
        out, ra, _ = oracle.invert(h, m, b, EVENODD, 1e-18)
        assert oracle.true_resid_full(h, m, out, b) < 2e-9
        assert ra.final_rsq < 1e-18
    oracle.free_links(h)


def test_invert_multi_matches_single(oracle):
    L = [4, 4, 4, 6]
    fat, lng = smeared_links(oracle, L, 51)
    h = oracle.links(L, fat, lng)
    b0, b1 = rand_vec(L, 52, EVEN), rand_vec(L, 53, EVEN)
    # flattened masses ascending: refine's m2[kx] index (asqtad_invert_p.c:299)
    # then equals m2[ko]; see test_invert_multi_refine_quirk for the other case
    masses = [[0.05, 0.1, 0.2], [0.3]]
    for eo in (EVEN, EVENODD):
        # solve_multi retargets the least-converged solve each outer
        # iteration, so several sources need more than the default 5 restarts
        outs, ras, _ = oracle.invert_multi(h, masses, [b0, b1], eo, 1e-20, max_restarts=50)
        k = 0
        for b, ms in zip((b0, b1), masses):
            for m in ms:
                single, _, _ = oracle.invert(h, m, b, EVEN, 1e-20)
                V = b.shape[0]
                d = np.linalg.norm(outs[k][: V // 2] - single[: V // 2]) / np.linalg.norm(single[: V // 2])
                assert d < 1e-7, (eo, m, d)
                tr = oracle.true_resid_even(h, m, outs[k], b)
                assert tr < 2e-8
                if eo == EVEN:  # reported residual is the true one, squared
                    assert ras[k].final_rsq == pytest.approx(tr * tr, rel=1e-3, abs=1e-26)
                k += 1
    # two-source trick path (nsrc=2, one mass each, different masses)
    outs, ras, _ = oracle.invert_multi(h, [[0.07], [0.15]], [b0, b1], EVENODD, 1e-20)
    assert oracle.true_resid_even(h, 0.07, outs[0], b0) < 1e-7
    assert oracle.true_resid_even(h, 0.15, outs[1], b1) < 1e-7
    oracle.free_links(h)


def test_relnorm2_closed_form(oracle):
    L = [4, 4, 4, 4]
    V = int(np.prod(L))
    h = oracle.links(L, *random_links(L, 61))
    x = np.ones((V, 3, 2))
    x[5] = 0  # a zero site contributes 1
    r = 2 * np.ones((V, 3, 2))
    got = oracle.relnorm2(h, r, x, EVEN)
    assert got == pytest.approx(((V // 2 - 1) * 4 + 1) / (V // 2))
    oracle.free_links(h)


def test_plaquette(oracle):
    L = [4, 4, 4, 4]
    V = int(np.prod(L))
    assert oracle.plaquette(L, free_field(L)) == pytest.approx(1.0)
    U = gauge_numpy(L, 71)
    p = oracle.plaquette(L, U)
    g = rand_su3(V, np.random.default_rng(72))
    assert oracle.plaquette(L, gauge_transform(L, U, g)) == pytest.approx(p, rel=1e-13)
    assert 0.5 < p < 1.0


def test_invert_multi_refine_quirk(oracle):
    """refine() uses m2[kx] (kx indexes the sorted unique masses) where it
    means m2[ko] (asqtad_invert_p.c:299).  With the flattened masses out of
    ascending order the Galerkin scale of those solves uses the wrong mass and
    solve_multi stalls; the oracle keeps this behaviour, and final_rsq reports
    the true (unconverged) residual."""
    L = [4, 4, 4, 6]
    fat, lng = smeared_links(oracle, L, 51)
    h = oracle.links(L, fat, lng)
    b0, b1 = rand_vec(L, 52, EVEN), rand_vec(L, 53, EVEN)
    outs, ras, _ = oracle.invert_multi(h, [[0.05, 0.1, 0.3], [0.2]], [b0, b1], EVEN, 1e-20,
                                       max_restarts=10)
    tr = [oracle.true_resid_even(h, m, o, b)
          for m, o, b in zip((0.05, 0.1, 0.3, 0.2), outs, (b0, b0, b0, b1))]
    assert tr[2] > 1e-3 and tr[3] > 1e-3  # the two swapped masses stall
    for k in (2, 3):
        assert ras[k].final_rsq == pytest.approx(tr[k] ** 2, rel=1e-6)
    oracle.free_links(h)


def test_dslash_dir_sums_to_hop(oracle):
    """The eight one-direction transports (4 directions x fb) with unit
    weights add up to the hop (style 1), which is also the independent
    check_op stencil at m = 0; each call accumulates into out."""
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    fat, lng = random_links(L, 81)
    h = oracle.links(L, fat, lng)
    v = rand_vec(L, 82)
    for eo_out in (EVEN, ODD):
        sl = slice(0, V // 2) if eo_out == EVEN else slice(V // 2, V)
        start = rand_vec(L, 83)
        out = start.copy()
        for mu in range(4):
            for fb in (1, -1):
                oracle.dslash_dir(h, mu, fb, 1.0, 1.0, v, eo_out, out)
        hop = oracle.dslash(h, 0.0, v, eo_out, 1 - eo_out)
        ref = oracle.check_op(L, fat, lng, 0.0, v)
        np.testing.assert_allclose(out[sl] - start[sl], hop[sl], rtol=0, atol=1e-12)
        np.testing.assert_allclose(out[sl] - start[sl], ref[sl], rtol=0, atol=1e-12)
        other = slice(V // 2, V) if eo_out == EVEN else slice(0, V // 2)
        assert np.array_equal(out[other], start[other])  # only eo_out is touched
    # weights: fat only (x2) equals twice the hop of the links without Naik
    h8 = oracle.links(L, fat, None)
    out = np.zeros((V, 3, 2))
    for mu in range(4):
        for fb in (1, -1):
            oracle.dslash_dir(h, mu, fb, 2.0, 0.0, v, EVEN, out)
    hop8 = oracle.dslash(h8, 0.0, v, EVEN, ODD)
    np.testing.assert_allclose(out[: V // 2], 2 * hop8[: V // 2], rtol=0, atol=1e-12)
    # style 0 sign (bsign -1): forward terms minus the stored backward ones
    fwd = np.zeros((V, 3, 2))
    for mu in range(4):
        oracle.dslash_dir(h, mu, 1, 1.0, 1.0, v, EVEN, fwd)
    s0 = np.zeros((V, 3, 2))
    for mu in range(4):
        for fb in (1, -1):
            oracle.dslash_dir(h, mu, fb, 1.0, 1.0, v, EVEN, s0, bsign=-1.0)
    hop = oracle.dslash(h, 0.0, v, EVEN, ODD)
    np.testing.assert_allclose(s0[: V // 2], 2 * fwd[: V // 2] - hop[: V // 2], rtol=0, atol=1e-12)
    oracle.free_links(h)
    oracle.free_links(h8)


def test_single_precision_mode(oracle):
    """FP32 storage mode (the F3 build): every stored value is FP32
    representable, the operator agrees with FP64 to FP32 accuracy, and an
    EVEN solve reaches the FP32-limited residual with the FP64 solution."""
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    fat, lng = smeared_links(oracle, L, 91)
    h64 = oracle.links(L, fat, lng)
    h32 = oracle.links(L, fat, lng)
    oracle.set_single(h32)
    v = rand_vec(L, 92).astype(np.float32).astype(np.float64)
    d64 = oracle.dslash(h64, 0.3, v, EVENODD, EVENODD)
    d32 = oracle.dslash(h32, 0.3, v, EVENODD, EVENODD)
    assert np.array_equal(d32, d32.astype(np.float32).astype(np.float64))
    err = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    assert 1e-9 < err < 1e-6
    b = rand_vec(L, 93, EVEN).astype(np.float32).astype(np.float64)
    x64, r64, _ = oracle.invert(h64, 0.1, b, EVEN, 1e-10)
    x32, r32, _ = oracle.invert(h32, 0.1, b, EVEN, 1e-10)
    assert np.array_equal(x32, x32.astype(np.float32).astype(np.float64))
    tr = oracle.true_resid_even(h64, 0.1, x32, b)
    assert tr < 1e-4
    d = np.linalg.norm(x32[: V // 2] - x64[: V // 2]) / np.linalg.norm(x64[: V // 2])
    assert d < 1e-4
    assert abs(r32.final_iter - r64.final_iter) <= max(3, 0.1 * r64.final_iter)
    oracle.free_links(h64)
    oracle.free_links(h32)


def test_mixed_rsq_branch(oracle):
    """solve_multi's FP32 inner branch (asqtad_invert_p.c:510-538): with a
    small mixed_rsq the first outer iteration runs in FP32 down to mixed_rsq,
    the FP64 refinement finishes; the result is the FP64 solution."""
    L = [4, 4, 4, 6]
    V = int(np.prod(L))
    fat, lng = smeared_links(oracle, L, 95)
    h = oracle.links(L, fat, lng)
    b = rand_vec(L, 96, EVEN)
    x64, r64, _ = oracle.invert(h, 0.1, b, EVEN, 1e-20)
    xm, rm, _ = oracle.invert(h, 0.1, b, EVEN, 1e-20, mixed_rsq=1e-8)
    assert rm.final_restart >= 1  # an FP32 outer iteration and at least one FP64 one
    assert oracle.true_resid_even(h, 0.1, xm, b) < 2e-9
    d = np.linalg.norm(xm[: V // 2] - x64[: V // 2]) / np.linalg.norm(x64[: V // 2])
    assert d < 1e-8
    # the FP32 stage is real: stopping it at 1e-8 costs more outer work than
    # the plain FP64 solve, which needs a single outer iteration
    assert r64.final_restart == 0
    oracle.free_links(h)
